"""CPU checker for the gjxl_amd HIP path.  TEST INFRASTRUCTURE ONLY.

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may use
this package; the product (gjxl_amd/) never imports it.
"""
