"""CPU oracle for the Projective Dynamics step (test infrastructure only).

Only tests/, ``__graft_entry__.smoke()`` and bench.py's ``cpu_baseline`` leg
may import this package, as the reference the GPU results are compared with
or as the CPU baseline; the product
package ``vistio_amd`` never does.
"""
