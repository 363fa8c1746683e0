"""ORACLE package -- test infrastructure only.

CPU restatements of the reference's projector / SIRT hot path, used as the
checker by tests/, by the smoke check and by bench.py's CPU-baseline leg.
The product package (vir_amd) never imports anything from here.
"""
