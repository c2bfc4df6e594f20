"""Entry point `python -m mammos_mumag_amd.scripts.hmag <name>` (see the package docstring)."""
import sys

from . import main_hmag

if __name__ == "__main__":
    main_hmag(sys.argv)
