"""DFSPH dam-break run (reference ``examples/dam_break/dfsph.py``)."""
from .run import main

if __name__ == "__main__":
    main("dfsph")
