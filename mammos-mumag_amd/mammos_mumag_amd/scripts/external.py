"""Entry point `python -m mammos_mumag_amd.scripts.external <name>` (see the package docstring)."""
import sys

from . import main_external

if __name__ == "__main__":
    main_external(sys.argv)
