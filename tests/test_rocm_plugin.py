"""The ROCm plugin's protocol and its native module's surface, on the CPU: what dynG's backend
selection asks of a plugin (python/dyng/_backend.py:1-30: available(), native, __version__), and
the names, signatures and enumeration values of dynG's extension module (python/dyng/_core.pyi)
that the sssp path of the typed layer reaches (python/dyng/{resources,graph,batch,sssp}.py). The
device calls themselves run in tests/test_gpu_rocm_plugin.py.
"""

from __future__ import annotations

import inspect
import sys
import types

import numpy as np
import pytest

import dyng_amd.rocm_plugin as plugin
from dyng_amd import errors

# name -> parameter names, as _core.pyi declares them (python/dyng/_core.pyi:91-147, 236-455,
# 1156-1390), for the callables the sssp path reaches
SIGNATURES = {
    "Resources.cuda": ("device", "stream", "host_threads"),
    "Resources.openmp": ("num_threads",),
    "GraphI32I32I32.from_edges": ("resources", "num_vertices", "src", "dst", "weights", "num_weights", "properties"),
    "GraphI32I32I32.from_csr": ("resources", "row_ptr", "col_ind", "weights", "num_weights", "properties"),
    "GraphI32I32I32.with_capacity": ("resources", "vertex_capacity", "row_capacity", "properties"),
    "GraphI32I32I32.apply": ("self", "resources", "batch"),
    "GraphI32I32I32.to_csr": ("self", "resources"),
    "GraphI32I32I32.clone": ("self", "resources"),
    "GraphI32I32I32.to_backend": ("self", "resources"),
    "GraphI32I32I32.reserve": ("self", "resources", "edge_capacity"),
    "EdgeBatchI32I32.__init__": ("self", "insert_src", "insert_dst", "insert_weights", "delete_src", "delete_dst",
                                 "num_weights", "insert_vertices", "insert_vertex_labels", "delete_vertices"),
    "SsspResultI32.set_options": ("self", "options"),
    "SsspResultI32.clone": ("self", "resources"),
    "sssp_compute": ("resources", "graph", "source", "options"),
    "sssp_update": ("resources", "graph", "batch", "result"),
    "sssp_from_arrays": ("resources", "graph", "source", "distances", "parents", "canonicalize", "options"),
}

ENUMS = {
    "Backend": {"sequential": 0, "openmp": 1, "cuda": 2},
    "MemorySpace": {"host": 0, "pinned_host": 1, "device": 2, "managed": 3},
    "CopyPolicy": {"allow": 0, "warn": 1, "error": 2},
    "Engine": {"automatic": 0, "fused": 1, "operators": 2},
    "RowLayout": {"compact": 0, "slotted": 1, "slack": 2},
    "VertexStatus": {"active": 0, "deleted": 1, "inactive": 2},
    "RowOrder": {"sorted": 0, "append": 1},
    "MultiEdges": {"forbid": 0, "allow": 1},
    "ExistingInsert": {"upsert": 0, "error": 1, "ignore": 2},
    "MissingDelete": {"ignore": 0, "error": 1},
    "SelfLoop": {"keep": 0, "drop": 1, "error": 2},
}

# attributes the typed layer reads of the native objects (python/dyng/{resources,graph,sssp}.py)
ATTRIBUTES = {
    "Resources": ("backend", "device", "stream", "num_threads", "default_space", "copy_policy", "workspace_bytes",
                  "release_workspaces", "warm_up", "synchronize"),
    "GraphI32I32I32": ("num_vertices", "num_active_vertices", "vertex_status", "num_edges", "num_weights", "directed",
                       "has_transposed", "space", "version", "properties", "to_csr", "apply", "clone", "to_backend",
                       "reserve", "check_integrity"),
    "SsspResultI32": ("source", "distances", "parents", "writer", "generation", "options", "set_options",
                      "graph_version", "space", "clone"),
}


def test_plugin_protocol():
    assert plugin.__version__ == "0.4.0"
    assert isinstance(plugin.available(), bool)
    st = plugin.status()
    assert plugin.available() == (st.reason == "")
    if not plugin.available():
        assert st.reason  # says why


def test_gpu_targets_are_read_from_the_driver_topology(tmp_path):
    for i, (cpu, target) in enumerate([(True, 0), (False, 90500), (False, 90402)]):
        d = tmp_path / str(i)
        d.mkdir()
        (d / "properties").write_text(f"cpu_cores_count {4 if cpu else 0}\ngfx_target_version {target}\n")
    assert plugin._gpu_targets(str(tmp_path)) == (90500, 90402)


def test_native_is_loaded_on_first_use_and_has_the_sssp_path():
    native = plugin.native
    for name in native.SSSP_PATH:
        assert hasattr(native, name), name
    assert native.__version__ == plugin.__version__
    assert native.build_config["cuda"] is True  # the selection requires a device module


def test_signatures_match_the_reference_bindings():
    native = plugin.native
    for path, params in SIGNATURES.items():
        obj = native
        for part in path.split("."):
            obj = getattr(obj, part)
        got = tuple(inspect.signature(obj).parameters)
        assert got == params, (path, got)


def test_enumerations_match_the_reference_values():
    native = plugin.native
    for name, values in ENUMS.items():
        assert {m.name: m.value for m in getattr(native, name)} == values, name


def test_objects_have_the_attributes_the_typed_layer_reads():
    native = plugin.native
    for cls, attrs in ATTRIBUTES.items():
        for a in attrs:
            assert hasattr(getattr(native, cls), a) or a in ("copy_policy", "writer", "generation"), (cls, a)
    r = native.Resources.__new__(native.Resources)
    r.copy_policy = native.CopyPolicy.warn
    assert r.copy_policy == native.CopyPolicy.warn


def test_options_semantics_and_properties_default_like_the_reference():
    native = plugin.native
    o = native.SsspOptions()
    assert (o.delta, o.objective, o.cuda_engine, o.validate_inputs) == (0, 0, native.Engine.automatic, True)
    s = native.BatchSemantics.upsert_last_wins()
    assert (s.on_existing_insert, s.on_missing_delete, s.on_self_loop) == (
        native.ExistingInsert.upsert, native.MissingDelete.ignore, native.SelfLoop.keep)
    assert native.BatchSemantics.set().as_sets and native.BatchSemantics.net_effect().as_net_effect
    p = native.GraphProperties.mosp_compatible()
    assert p.order == native.RowOrder.append and p.parallel_edges == native.MultiEdges.allow
    api = p._to_api()
    assert api.order == "append" and api.parallel_edges == "allow" and api.store_transposed
    e = native.GraphProperties()
    e.semantics.on_existing_insert = native.ExistingInsert.error
    assert e._to_api().semantics.on_existing_insert == "error"
    with pytest.raises(errors.NotSupportedError):
        bad = native.GraphProperties()
        bad.semantics = native.BatchSemantics.set()
        bad._to_api()


def test_edge_batch_arrays():
    native = plugin.native
    i32 = np.int32
    b = native.EdgeBatchI32I32(np.array([0, 1], i32), np.array([2, 3], i32), np.array([5, 6, 7, 8], i32),
                               np.array([4], i32), np.array([5], i32), 2, None, None, None)
    assert (b.num_insertions, b.num_deletions, b.num_weights) == (2, 1, 2)
    assert b._w.tolist() == [[5, 6], [7, 8]]  # edge-major weights, num_weights per insertion
    with pytest.raises(TypeError):
        native.EdgeBatchI32I32(np.array([0], np.int64), np.array([1], i32), None, np.array([], i32),
                               np.array([], i32), 1, None, None, None)
    with pytest.raises(errors.NotSupportedError):
        native.EdgeBatchI32I32(np.array([], i32), np.array([], i32), None, np.array([], i32), np.array([], i32), 1,
                               np.array([3], i32), None, None)


def test_names_off_the_path_say_so():
    native = plugin.native
    with pytest.raises(errors.NotSupportedError, match="not on the HIP path"):
        native.cycle_count_compute
    with pytest.raises(errors.NotSupportedError):
        native.Resources.openmp(4)


def test_errors_are_raised_as_the_reference_classes(monkeypatch):
    native = plugin.native

    class InvalidArgumentError(ValueError):
        pass

    fake = types.ModuleType("dyng.errors")
    fake.InvalidArgumentError = InvalidArgumentError
    monkeypatch.setitem(sys.modules, "dyng", types.ModuleType("dyng"))
    monkeypatch.setitem(sys.modules, "dyng.errors", fake)

    @native._translated
    def fails():
        raise errors.InvalidArgumentError("sssp: the source 9 is out of range [0, 4)")

    with pytest.raises(InvalidArgumentError, match="out of range"):
        fails()
    monkeypatch.delitem(sys.modules, "dyng.errors")
    monkeypatch.delitem(sys.modules, "dyng")
    with pytest.raises(errors.InvalidArgumentError):
        fails()
