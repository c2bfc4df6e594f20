"""Entry point `python -m mammos_mumag_amd.scripts.exani <name>` (see the package docstring)."""
import sys

from . import main_exani

if __name__ == "__main__":
    main_exani(sys.argv)
