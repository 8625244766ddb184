"""WCSPH dam-break run (reference ``examples/dam_break/wcsph.py``)."""
from .run import main

if __name__ == "__main__":
    main("wcsph")
