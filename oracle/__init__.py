"""ORACLE -- test infrastructure only.

CPU restatements of the reference's hot path (strict-IEEE float for the
voxelizer, float64 transcriptions of the shaders).  Only tests/,
__graft_entry__.smoke() and bench.py's CPU-baseline leg may use this package,
and only as the checker; the product library never links or calls it.
"""
