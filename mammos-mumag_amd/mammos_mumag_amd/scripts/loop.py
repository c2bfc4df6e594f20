"""Entry point `python -m mammos_mumag_amd.scripts.loop <name>` (see the package docstring)."""
import sys

from . import main_loop

if __name__ == "__main__":
    main_loop(sys.argv)
