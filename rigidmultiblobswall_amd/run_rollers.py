"""Former name of the simulation driver, kept for callers that import it:
rigidmultiblobswall_amd.run_simulation runs rollers and rigid bodies."""
from .run_simulation import blob_blob_name, build, main, mobility_precision, run  # noqa: F401

if __name__ == "__main__":
    import sys
    sys.exit(main())
