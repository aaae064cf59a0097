"""Shared helpers of the GPU parity tests: run the HIP path and compare it with
the CPU oracle on the same inputs."""
import sys

import numpy as np


def announce(*what):
    print("CASE", *what, flush=True)
    sys.stdout.flush()


def expected_canonical(oracle, A_np, R_np):
    want = A_np.copy()
    st = oracle.canonical(want, R_np)
    assert st == oracle.OK, f"oracle rejected the input: {st}"
    return want


def first_mismatch(a, b):
    if a.shape[0] == 0:
        return -1
    neq = np.nonzero(np.any((a != b).reshape(a.shape[0], -1), axis=1))[0]
    return int(neq[0]) if neq.size else -1


def gpu_remove_parity(ctx, oracle, A, R, config=None, flags=0, want=None):
    """A: CPU torch tensor (n,) or (n, w) int32; R: CPU torch int32 (k,).
    Runs ctx.remove on copies on the device and checks the survivors are
    array-identical to the canonical oracle and satisfy every invariant."""
    import torch

    n, k = A.shape[0], R.shape[0]
    A_np = A.numpy()
    R_np = np.ascontiguousarray(R.numpy())
    if want is None:
        want = expected_canonical(oracle, A_np, R_np)
    dA = A.cuda()
    dR = R.cuda()
    new_len = ctx.remove(dA, dR, config=config, flags=flags)
    torch.cuda.synchronize()
    assert ctx.check_errors() == 0  # raises RedzoneError(RZ_EDEVICE) if a device-side check failed
    got = dA.cpu().numpy()
    assert new_len == n - k
    m = first_mismatch(got[: n - k], want[: n - k])
    assert m < 0, f"first mismatch at {m}: got {got[m]} want {want[m]} (n={n}, k={k}, cfg={config})"
    bad, fb = oracle.check(A_np, got, R_np)
    assert bad == 0, f"invariant violation {bad} at {fb}"
    return got, dR
