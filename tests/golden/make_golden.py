"""Write the golden fixtures of tests/test_golden.py from the real reference.

Run in a container that holds the reference and has built its driver
(``make -C oracle ref``): every expected output here comes from the
reference's own cpu::single_thread free functions, called through
oracle/_ref/librpp_ref.so on seeded inputs.  The fixtures (inputs and outputs,
FP64, small batches at the shapes of the benchmark workloads) are committed,
so the parity tests can use them where the reference is absent.

    python3 tests/golden/make_golden.py [fixture names: only those]
"""
import hashlib
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "..", "..", "oracle"))

import oracle as O  # noqa: E402


def dim_of(w, n):
    return O.tensor_dimension(w, n)


def damped(rng, n, w, N):
    db = [0]
    p = 1
    for _ in range(N + 1):
        db.append(db[-1] + p)
        p *= w
    x = rng.uniform(-1.0, 1.0, size=(n, db[-1])) * min(1.0, 16.0 / np.sqrt(db[-1]))
    for k in range(N + 1):
        x[:, db[k]:db[k + 1]] *= 0.5 ** k
    return x


def main():
    if not O.has_ref():
        raise SystemExit("the reference driver (oracle/_ref) is not built here")
    R = O.refimpl
    out = {}
    # products at the (4,4) shape: 16 pairs, U[-1,1], degree 0 = 1, full windows
    rng = np.random.default_rng(0xC1)
    w, N = 4, 4
    dim = dim_of(w, N)
    lhs = rng.uniform(-1, 1, (16, dim))
    rhs = rng.uniform(-1, 1, (16, dim))
    lhs[:, 0] = 1.0
    rhs[:, 0] = 1.0
    o = np.zeros((16, dim))
    R.ft_mul(w, N, o, lhs, rhs)
    out["mul_w4d4"] = dict(lhs=lhs, rhs=rhs, out=o)
    # the same with windows out (1,3), lhs (0,2), rhs (1,4) and beta = 0.5
    o = np.zeros((16, dim))
    R.ft_mul(w, N, o, lhs, rhs, ow=(1, 3), lw=(0, 2), rw=(1, 4), beta=0.5)
    out["mul_w4d4_windows"] = dict(lhs=lhs, rhs=rhs, out=o)
    # fma, the three in-place variants: a, b, c at (3,4)
    rng = np.random.default_rng(31)
    w, N = 3, 4
    dim = dim_of(w, N)
    a, b, c = (damped(rng, 4, w, N) for _ in range(3))
    o = np.zeros((4, dim))
    R.ft_fma(w, N, o, a, b, c, alpha=0.75, beta=1.25)
    out["fma_w3d4"] = dict(a=a, b=b, c=c, out=o)
    for variant in (231, 123, 213):
        x = a.copy()
        R.ft_inplace_fma(variant, w, N, x, b, c, alpha=0.5, beta=2.0)
        out[f"inplace_fma{variant}_w3d4"] = dict(a=a, b=b, c=c, out=x)
    # exp, log, antipode, reflect, adjoints at (3,4)
    for name in ("ft_exp", "ft_log"):
        o = np.zeros((4, dim))
        getattr(R, name)(w, N, o, a)
        out[f"{name[3:]}_w3d4"] = dict(arg=a, out=o)
    for name in ("tensor_antipode", "tensor_reflect"):
        o = np.zeros((4, dim))
        getattr(R, name)(w, N, o, a)
        out[f"{name[7:]}_w3d4"] = dict(arg=a, out=o)
    for name in ("ft_adj_lmul", "ft_adj_rmul"):
        o = np.zeros((4, dim))
        getattr(R, name)(w, N, o, a, b)
        out[f"{name[3:]}_w3d4"] = dict(op=a, arg=b, out=o)
    # pairing at (6,4): 4 objects against their functionals
    rng = np.random.default_rng(0xC5F)
    w, N = 6, 4
    dim = dim_of(w, N)
    f = rng.normal(size=(4, dim))
    x = damped(rng, 4, w, N)
    o = np.zeros(4)
    R.tensor_pairing(w, N, o, f, x)
    out["pairing_w6d4"] = dict(functional=f, arg=x, out=o)
    # the chained ft_fmexp of the headline workload: 2 paths x 16 increments at (5,6)
    rng = np.random.default_rng(0xC2)
    inc = rng.normal(0.0, np.sqrt(1.0 / 128), size=(2, 16, 5))
    out["signature_w5d6"] = dict(increments=inc, out=O.ref_signature(5, 6, inc))
    # the log-signature pipeline at (3,8): 2 paths x 64 increments, then ft_log,
    # tensor_to_lie (T2L) and lie_to_tensor (L2T) through the reference ops
    rng = np.random.default_rng(0xC3)
    inc = rng.normal(0.0, np.sqrt(1.0 / 64), size=(2, 64, 3))
    sig = O.ref_signature(3, 8, inc)
    lg = np.zeros_like(sig)
    R.ft_log(3, 8, lg, sig)
    lie = O.ref_lie_convert(True, True, 3, 8, lg)
    back = O.ref_lie_convert(False, True, 3, 8, lie)
    out["logsig_w3d8"] = dict(increments=inc, signature=sig, log=lg, lie=lie, back=back)
    # ft_log on the worst-conditioned log-signatures of the width 3, depth 8
    # workload (16384 paths x 64 increments N(0, 1/64), seed 0xC3): the 32
    # paths whose log-signature has the largest coefficient (up to ~30, where
    # the Horner sums cancel most).  Kept: the increments, the reference's
    # ft_log of the reference's signatures, and a SHA-256 of those signatures
    # (the oracle's chain rebuilds them bit for bit, which the test checks)
    n_all, L = 16384, 64
    rng = np.random.default_rng(0xC3)
    inc_all = rng.normal(0.0, np.sqrt(1.0 / L), size=(n_all, L, 3))
    sig_all = O.ref_signature(3, 8, inc_all, threads=os.cpu_count() or 1)
    lg_all = np.zeros_like(sig_all)
    R.ft_log(3, 8, lg_all, sig_all)
    worst = np.sort(np.argsort(-np.abs(lg_all).max(axis=1), kind="stable")[:32])
    sig_w = np.ascontiguousarray(sig_all[worst])
    out["log_c3_worst_w3d8"] = dict(paths=worst.astype(np.int64), increments=inc_all[worst],
                                    log=lg_all[worst],
                                    signature_sha256=np.frombuffer(hashlib.sha256(sig_w.tobytes()).digest(),
                                                                   dtype=np.uint8))
    # linear operations on partial windows, tensor (3,4) and Lie (3,4)
    # gradings: 3 objects each
    rng = np.random.default_rng(0x11)
    for kind, start in (("tensor", O.degree_begin(3, 4)), ("lie", O.lie_degree_begin(3, 4))):
        dim = start[-1]
        x, y, z = (rng.uniform(-1, 1, (3, dim)) for _ in range(3))
        assign = np.full((3, dim), 7.0)
        R.vector_assign(start, 4, assign, y, ow=(1, 3), aw=(2, 4))
        add = np.full((3, dim), 7.0)
        R.vector_add(start, 4, add, x, y, ow=(0, 3), lw=(1, 4), rw=(0, 2), alpha=0.75, beta=-1.5)
        inplace = z.copy()
        R.vector_inplace_add(start, 4, inplace, x, lw=(1, 4), rw=(0, 2), alpha=-0.25)
        scaled = z.copy()
        R.vector_scalar_multiply(start, 4, scaled, 1.25, vw=(2, 4))
        const = z.copy()
        R.vector_set_constant(start, 4, const, -0.5, vw=(0, 1))
        ident = z.copy()
        R.tensor_set_identity(start, 4, ident, 2.0, vw=(0, 2))
        addid = z.copy()
        R.tensor_add_identity(start, 4, addid, 0.5)
        out[f"linear_{kind}_w3d4"] = dict(x=x, y=y, z=z, assign=assign, add=add, inplace=inplace, scaled=scaled,
                                          const=const, ident=ident, addid=addid)
    # the shuffle family at (3,4): 3 objects, damped inputs, partial windows
    rng = np.random.default_rng(0x5F)
    w, N = 3, 4
    dim = dim_of(w, N)
    x, y, z = (damped(rng, 3, w, N) for _ in range(3))
    sh = dict(x=x, y=y, z=z)
    o = np.full((3, dim), 7.0)
    R.st_mul(w, N, o, x, y, ow=(0, 4), lw=(0, 3), rw=(1, 4), beta=-0.5)
    sh["st_mul"] = o
    o = np.full((3, dim), 7.0)
    R.st_fma(w, N, o, z, x, y, ow=(1, 4), aw=(0, 2), bw=(0, 4), cw=(0, 4), alpha=0.75, beta=1.25)
    sh["st_fma"] = o
    o = z.copy()
    R.st_inplace_fma(w, N, o, x, y, aw=(0, 3), bw=(1, 4), cw=(0, 4), alpha=0.5, beta=2.0)
    sh["st_inplace_fma"] = o
    o = np.full((3, dim), 7.0)
    R.st_adj_mul(w, N, o, x, y, ow=(0, 4), pw=(0, 2), gw=(0, 4))
    sh["st_adj_mul"] = o
    o = np.full((3, dim), 7.0)
    R.left_hs_mul(w, N, o, x, y, ow=(0, 4), xw=(1, 4), yw=(0, 3), beta=1.5)
    sh["left_hs_mul"] = o
    for name in ("left_hs_adj_lmul", "left_hs_adj_rmul"):
        o = np.full((3, dim), 7.0)
        getattr(R, name)(w, N, o, x, y, ow=(0, 3), xw=(0, 4), yw=(0, 4))
        sh[name] = o
    out["shuffle_w3d4"] = sh
    only = set(sys.argv[1:])
    for name, arrays in out.items():
        if only and name not in only:
            continue
        np.savez_compressed(os.path.join(HERE, name + ".npz"), **arrays)
    print("wrote", len(only) if only else len(out), "fixtures to", HERE)


if __name__ == "__main__":
    main()
