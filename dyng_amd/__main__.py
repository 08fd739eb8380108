"""``python -m dyng_amd``: the command line (dyng_amd.cli)."""

import sys

from .cli import main

sys.exit(main())
