"""Python-side common layer, mirroring sqaodpy/sqaod/common."""
