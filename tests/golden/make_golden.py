"""Writes the golden fixtures under tests/golden/ from the reference's own tests.

Run in a container that has the reference checkout at /root/reference:
    python tests/golden/make_golden.py
The outputs are data only: the expected values the reference's unit tests
assert (transcribed with their file:line), and the PLINK fixture files the
reference's tests hold (data/testdata/ld.{bed,bim,fam}, copied byte for byte).
"""
import json
import os
import shutil

HERE = os.path.dirname(os.path.abspath(__file__))
REF = "/root/reference"

golden = {
    # score/batch.rs:1012-1036 -- one byte 0b01_11_10_00 = people
    # (p0=00, p1=10, p2=11, p3=01), W_eff 1.5, corr 2.0; variant-major partial
    # scores (before the baseline) and missing counts.
    "variant_major_missing_correction": {
        "byte": 0b01_11_10_00, "num_people": 4, "w_eff": 1.5, "corr": 2.0,
        "partial_scores": [0.0, 1.5, 3.0, -2.0], "missing": [0, 0, 0, 1],
    },
    # score/batch.rs:1038-1066 -- 101 people, genotype codes set for four
    # people (person, code), everything else 00.
    "word_skip_and_tail": {
        "num_people": 101, "w_eff": 1.5, "corr": 2.0,
        "codes": [[3, 2], [32, 3], [99, 1], [100, 2]],
        "partial_scores": {"3": 1.5, "32": 3.0, "99": -2.0, "100": 1.5},
        "missing": {"99": 1}, "nonzero_scores": 4, "missing_total": 1,
    },
    # score/batch.rs:1068-1114 -- 577 scores with weight 1.0, three people
    # with dosages 1, 2, 0.
    "many_scores": {"num_people": 3, "num_scores": 577, "dosages": [1, 2, 0],
                    "scores": [1.0, 2.0, 0.0]},
    # score/prepare.rs:1040-1097 -- duplicate aggregation in A2-dosage space.
    "aggregation": {
        "same_orientation": {"rows": [[0.25, False], [-0.10, False]], "w_eff": 0.15, "corr": 0.0},
        "flipped": {"rows": [[0.40, True]], "w_eff": -0.40, "corr": 0.80},
        "mixed_rows": [[0.35, False], [0.10, True], [-0.05, False]],
        "tolerance": 1e-6,
    },
    # score/tests/sim_test.py:202-239 -- id_special_case: effect allele A is
    # BIM A1 (flipped), weight 0.1 on 50 loci; A2 dosages 2 (x10), 1 (x30),
    # 0 (x10) -> effect-allele sum 5.0 over 50 scored loci, AVG 0.1.
    "special_case": {"weight": 0.1, "a2_dosages": [2] * 10 + [1] * 30 + [0] * 10,
                     "flipped": True, "sum": 5.0, "avg": 0.1},
    # score/tests/gpu_tests.rs:611-628 -- deterministic genotype generator and
    # 2-bit encoding of the GPU parity tests.
    "gpu_tests_generator": {
        "key": "(person*131 + variant*17 + 23) % 31",
        "classes": {"missing": [0, 1], "dosage0": [2, 9], "dosage1": [10, 20], "dosage2": [21, 30]},
        "codes": {"dosage0": 0, "missing": 1, "dosage1": 2, "dosage2": 3},
    },
    # score/cuda_backend.rs:2432-2438 -- split self-test inputs (64^3):
    # dosage[i] = i % 3 over [64 people][64 variants], weight[i] =
    # ((7 i mod 101) - 50) / 13 over [64 variants][64 scores].
    "split_self_test": {"dim": 64},
    # score/tests/gpu_tests.rs:11-15, 134-162 -- GPU/CPU agreement rule.
    "tolerance_rule": {"abs": 1e-6, "rel": 1e-6, "rel_floor": 1.0},
    # score/cuda_backend.rs:38-39, 608-617 -- device size gate.
    "size_gate": {"min_output_cells": 100000, "min_variant_score_work": 100000000},
}

with open(os.path.join(HERE, "golden.json"), "w") as f:
    json.dump(golden, f, indent=1, sort_keys=True)
for ext in ("bed", "bim", "fam"):
    shutil.copyfile(os.path.join(REF, "data", "testdata", "ld." + ext), os.path.join(HERE, "ld." + ext))
print("wrote", HERE)
