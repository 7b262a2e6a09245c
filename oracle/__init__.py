"""CPU oracle of the collide-stream step -- test infrastructure only.

Only the test suite, the smoke check and the CPU-baseline leg of the
benchmark may import this package; the shipped library never does.
"""
