"""Entry point `python -m mammos_mumag_amd.scripts.mapping <name>` (see the package docstring)."""
import sys

from . import main_mapping

if __name__ == "__main__":
    main_mapping(sys.argv)
