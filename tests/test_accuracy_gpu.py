"""Held-out accuracy of GPU training against the reference-shaped CPU trainer.

The engine's random stream is its own (the reference's RNG lives in an
unvendored dependency), so training trajectories differ from the reference's
and parity is statistical: mean held-out accuracy over 5 seeds.  The CPU side
is the oracle's reference-shaped trainer (per-thread PCG32 streams and
Fisher-Yates shuffles, OpenMP static over chunks, evaluation before the
negative-sampling test, as capi.c:1112-1191).

Workload 1: 8192 train / 2048 test samples, 256 Bernoulli(0.5) features,
class = 2*(x3 AND x17) + (x40 XOR x91) with 5% flips; 256 clauses/class,
target 64, specificity 8, 10 epochs.
"""
import numpy as np
import pytest

import santoku.tsetlin as tm
from oracle import oracle as O
from santoku.tsetlin import data as D

pytestmark = pytest.mark.gpu

W1 = dict(classes=4, features=256, clauses=256, target=64, specificity=8.0)


def accuracies(o, seeds, epochs=10):
    gpu, cpu = [], []
    for seed in seeds:
        d = D.xor_and_task(seed)
        eng = tm.classifier(dict(o, seed=seed))
        eng.train(samples=8192, problems=d["train_rows"], solutions=d["train_y"], iterations=epochs)
        gpu.append(float((eng.predict(d["test_rows"], 2048) == d["test_y"]).mean()))
        ref = O.Machine(O.params("classifier", **dict(o, seed=seed)))
        # 8 threads: the reference-shaped trajectory depends on the thread count
        ref.train_refshape(d["train_rows"], labels=d["train_y"], epochs=epochs, threads=8)
        cpu.append(float((ref.predict(d["test_rows"]) == d["test_y"]).mean()))
    return np.array(gpu), np.array(cpu)


def test_workload1_accuracy_within_half_point():
    """The pinned configuration (clause tolerance = maximum = 64, the spec
    defaults): mean held-out accuracy over 5 seeds within +-0.5 pt."""
    gpu, cpu = accuracies(dict(W1, clause_tolerance=64, clause_maximum=64), range(5))
    assert abs(gpu.mean() - cpu.mean()) <= 0.005, (gpu, cpu)


def test_learning_setting_agrees_with_reference_shape():
    """A setting where both trainers learn past the majority class (0.37):
    tolerance = maximum = 3.  Per-seed spread is about 2 pt here, so the 5-seed
    means are held to 1.5 pt (a CPU sweep of six such settings gave
    differences from -1.2 to +0.5 pt, of both signs)."""
    gpu, cpu = accuracies(dict(W1, clause_tolerance=3, clause_maximum=3, specificity=8.0, target=16), range(5))
    assert gpu.mean() > 0.45 and cpu.mean() > 0.45, (gpu, cpu)
    assert abs(gpu.mean() - cpu.mean()) <= 0.015, (gpu, cpu)


def test_encoder_hamming_within_half_point():
    """Workload 3's machine settings (tolerance = maximum = 64, specificity
    20, target 128) at a size the CPU trains in seconds: 256 visible bits ->
    16 hidden bits of the Gaussian-projection task, 128 clauses per bit,
    4096 train / 1024 held-out samples, 5 epochs.  Mean per-bit Hamming error
    over 5 seeds within +-0.5 pt of the reference-shaped trainer."""
    gpu, cpu = [], []
    for seed in range(5):
        d = D.projection_codes_task(seed, n=5120, visible=256, hidden=16)
        tr, te = slice(0, 4096), slice(4096, 5120)
        o = dict(hidden=16, visible=256, clauses=128, clause_tolerance=64, clause_maximum=64, specificity=20.0,
                 target=128, seed=seed)
        eng = tm.encoder(o)
        eng.train(samples=4096, sentences=d["rows"][tr], codes=d["codes"][tr], iterations=5)
        got = eng.predict(d["rows"][te], 1024)
        gpu.append(float(np.unpackbits(got ^ d["codes"][te], axis=1, bitorder="little")[:, :16].mean()))
        ref = O.Machine(O.params("encoder", classes=16, features=256, clauses=128, clause_tolerance=64,
                                 clause_maximum=64, specificity=20.0, target=128, seed=seed))
        ref.train_refshape(d["rows"][tr], codes=d["codes"][tr], epochs=5, threads=8)
        want = ref.predict(d["rows"][te])
        cpu.append(float(np.unpackbits(want ^ d["codes"][te], axis=1, bitorder="little")[:, :16].mean()))
    gpu, cpu = np.array(gpu), np.array(cpu)
    assert gpu.mean() < 0.25 and cpu.mean() < 0.25, (gpu, cpu)      # both learn (chance is 0.5)
    assert abs(gpu.mean() - cpu.mean()) <= 0.005, (gpu, cpu)
