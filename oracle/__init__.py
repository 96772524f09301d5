"""CPU oracle of the polygenic-scoring hot path -- TEST INFRASTRUCTURE ONLY.

Importable by tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg;
never by the product package gnomon_amd.
"""
