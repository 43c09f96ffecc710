"""CPU oracle of the D2Q9 BGK time step -- test infrastructure only.

Only tests, the smoke check and the CPU-baseline leg of bench.py use this
package, and only as the checker; the library (portlbm_amd) never does.
"""
