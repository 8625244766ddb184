"""PBF dam-break run (reference ``examples/dam_break/pbf.py``)."""
from .run import main

if __name__ == "__main__":
    main("pbf")
