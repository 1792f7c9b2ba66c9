"""GPU parity of the drop-in path: the C++ facade (include/dnn_solver.h) over
libdnnsolver.so, compiled and run the way a walker uses it
(tests/dropin_check.cpp, the INTEGRATION.md section-1 program).

Fixed batches (DNNSolverPoisson / DNNSolverGrad::sample on host vectors after
setSeed, dnn_solver_poisson.cpp:35-63, dnn_solver_gradient.cpp:40-73): every
hop, W*sign and eps_centre equals the oracle's draw (cdf_draw / locate on the
query's Philox uniforms), bit for bit, given
  network backend: the GPU's own kernels (drw_solver_eval_kernels on the same
                   descriptions and weights);
  FDM backend:     the GPU's FDM kernels G (drw_fdm_solve at the same
                   tolerance) of the oracle-voxelized cubes.
ParallelSampler (parallel_sampler.cpp:42-211) with 8 submitting threads in
front of the same solvers: each future's result is a valid draw for its own
query, the solver consumed exactly the counters [c1, c1 + n), and there is a
one-to-one assignment of those counters to the queries that reproduces every
result (a perfect matching); on the FDM backend eps_centre and |W| equal the
fixed batch's.
"""
import os
import shutil
import struct
import subprocess

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
B = 256
SEED, C0, C1 = 0x0DDC0FFEE, 5000, 90000
FDM_TOL = 1e-6
GRAD_CODE = {0: 2, 1: 3, 2: 1}  # axis -> WVTX, WVTY, WVTZ
AXIS_FACE = {0: {0: 4, 1: 5, 4: 1, 5: 0}, 1: {0: 2, 1: 3, 2: 1, 3: 0}, 2: {}}  # cuda_kernels.cu:146-153


@pytest.fixture(scope="module")
def run(tmp_path_factory):
    import torch  # noqa: F401  (one HIP runtime in this process)
    from deeprwcap_amd import _lib, structures as S
    assert shutil.which("g++"), "needs g++"
    tmp = tmp_path_factory.mktemp("dropin")
    libdir = os.path.dirname(_lib.LIB_PATH)
    exe = str(tmp / "dropin_check")
    subprocess.run(["g++", "-std=c++17", "-O1", "-pthread", f"-I{os.path.join(ROOT, 'include')}",
                    os.path.join(ROOT, "tests", "dropin_check.cpp"), f"-L{libdir}", "-ldnnsolver",
                    f"-Wl,-rpath,{libdir}", "-o", exe], check=True, timeout=300)
    L, C, _ = S.layered_boxes(B, 23, seed=314)
    axis = (np.arange(B) * 7 % 3).astype(np.int32)
    inp, outp = str(tmp / "in.bin"), str(tmp / "out.bin")
    with open(inp, "wb") as f:
        f.write(struct.pack("<iQQQ", B, SEED, C0, C1))
        f.write(L.astype(np.float32).tobytes())
        f.write(C.astype(np.float32).tobytes())
        f.write(axis.tobytes())
    print(f"dropin_check B {B} (models dir cwd)", flush=True)
    r = subprocess.run([exe, inp, outp, repr(FDM_TOL)], cwd=os.path.join(ROOT, "models"), capture_output=True,
                       text=True, timeout=240)
    print(r.stdout, r.stderr, flush=True)
    assert r.returncode == 0, r.stderr
    raw = open(outp, "rb").read()
    widths = (3, 5, 3, 5)
    vals = np.frombuffer(raw[:-32], dtype=np.float32)
    fixed, par, o = [], [], 0
    for w in widths:
        fixed.append(vals[o:o + B * w].reshape(B, w))
        o += B * w
    for w in widths:
        par.append(vals[o:o + B * w].reshape(B, w))
        o += B * w
    assert o == vals.size
    counters = np.frombuffer(raw[-32:], dtype=np.uint64)
    return dict(L=L, C=C, axis=axis, fixed=fixed, par=par, counters=counters)


@pytest.fixture(scope="module")
def nn_kernels(run):
    """Probabilities, W and all six kernels per query from the GPU networks
    (shipped weights), per solver kind."""
    import torch
    from deeprwcap_amd import _lib
    from deeprwcap_amd.dnnsolver import DNNSolverGrad, DNNSolverPoisson
    out = {}
    dL, dC = torch.from_numpy(run["L"]).cuda(), torch.from_numpy(run["C"]).cuda()
    dA = torch.from_numpy(run["axis"]).cuda()
    for kind, cls in ((0, DNNSolverPoisson), (1, DNNSolverGrad)):
        s = cls(0)
        s.loadModels(models_txt=os.path.join(ROOT, "models", "models.txt"))
        probs = torch.empty((B, 6), device="cuda")
        wts = torch.empty(B, device="cuda")
        kern = torch.empty((B, 6, 529), device="cuda")
        _lib.check(_lib.lib().drw_solver_eval_kernels(s.handle, B, dL.data_ptr(), dC.data_ptr(),
                                                       dA.data_ptr() if kind else None, probs.data_ptr(),
                                                       wts.data_ptr(), kern.data_ptr(), None), "eval_kernels")
        torch.cuda.synchronize()
        out[kind] = (probs.cpu().numpy(), wts.cpu().numpy(), kern.cpu().numpy())
    return out


@pytest.fixture(scope="module")
def fdm_kernels(run, orc):
    """GPU FDM kernels of the oracle-voxelized cubes (GFT, and the axis' code)."""
    from deeprwcap_amd import device
    cubes, _ = orc.voxelize(run["L"], run["C"], 23, normalise=False)
    res = {}
    for kind in (0, 1):
        codes = np.array([0 if kind == 0 else GRAD_CODE[int(a)] for a in run["axis"]], np.int32)
        pk = device.PackedCubes([23] * B)
        pk.set_cubes(list(cubes))
        it, rr = device.fdm_solve(pk, codes, tol=FDM_TOL)
        assert (it.cpu().numpy() > 0).all() and (rr.cpu().numpy() <= FDM_TOL).all()
        G = pk.G.cpu().numpy().reshape(B, 6 * 529)
        res[kind] = G
    cen = np.array([orc.center_eps(cubes[b].astype(np.float64), GRAD_CODE[int(run["axis"][b])]) for b in range(B)],
                   np.float32)
    return res, cubes, cen


def _nn_draw(orc, kind, k, b, u, axis):
    probs, wts, kern = k
    f = orc.cdf_draw(probs[b], int(u[0]))
    t = orc.cdf_draw(kern[b, f], int(u[1]))
    if kind == 0:
        return np.concatenate([orc.locate(23, t, f)])
    fo = AXIS_FACE[int(axis[b])].get(f, f)
    return np.concatenate([orc.locate(23, t, fo), [np.float32(wts[b] * np.sign(kern[b, f, t]))]])


def _fdm_draw(orc, G, b, u):
    i = orc.cdf_draw(G[b], int(u[0]))
    f, t = divmod(i, 529)
    return orc.locate(23, t, f), i


def test_fixed_batch_network_hops(run, nn_kernels, orc):
    cubes, mx = orc.voxelize(run["L"], run["C"], 23, normalise=True)
    for kind in (0, 1):
        out = run["fixed"][kind]
        bad = []
        for b in range(B):
            want = _nn_draw(orc, kind, nn_kernels[kind], b, orc.query_uniforms(SEED, C0 + b), run["axis"])
            w = len(want)
            if not np.array_equal(out[b, :w], want):
                bad.append((b, out[b, :w].tolist(), want.tolist()))
            if kind == 1 and out[b, 4] != np.float32(cubes[b, 11, 11, 11] * mx[b]):
                bad.append((b, "eps_centre", float(out[b, 4])))
        assert not bad, f"kind {kind}: {len(bad)} of {B} differ, first {bad[:3]}"


def test_fixed_batch_fdm_hops(run, fdm_kernels, orc):
    G, _, cen = fdm_kernels
    for kind in (0, 1):
        out = run["fixed"][2 + kind]
        for b in range(B):
            p, i = _fdm_draw(orc, G[kind], b, orc.query_uniforms(SEED, C0 + b))
            np.testing.assert_array_equal(out[b, :3], p, err_msg=f"kind {kind} query {b}")
            if kind == 1:
                w = np.abs(G[1][b]).sum() * np.sign(G[1][b][i])
                assert out[b, 3] == pytest.approx(w, rel=1e-6), (b, out[b, 3], w)
                assert out[b, 4] == cen[b]


def _perfect_matching(cands):
    """Kuhn's augmenting paths: does every query get its own counter?"""
    owner = {}

    def augment(q, seen):
        for c in cands[q]:
            if c in seen:
                continue
            seen.add(c)
            if c not in owner or augment(owner[c], seen):
                owner[c] = q
                return True
        return False

    return sum(augment(q, set()) for q in range(len(cands)))


@pytest.mark.parametrize("k", [0, 1, 2, 3], ids=["nn_poisson", "nn_gradient", "fdm_poisson", "fdm_gradient"])
def test_parallel_sampler_results(k, run, nn_kernels, fdm_kernels, orc):
    assert int(run["counters"][k]) == C1 + B, "the solver must consume exactly the counters [c1, c1 + n)"
    par = run["par"][k]
    kind = k % 2
    G, _, cen = fdm_kernels
    U = [orc.query_uniforms(SEED, C1 + j) for j in range(B)]
    cands = []
    for b in range(B):
        ok = []
        for j in range(B):
            if k < 2:
                want = _nn_draw(orc, kind, nn_kernels[kind], b, U[j], run["axis"])
                hit = np.array_equal(par[b, :len(want)], want)
            else:
                p, i = _fdm_draw(orc, G[kind], b, U[j])
                hit = np.array_equal(par[b, :3], p) and (kind == 0 or np.sign(par[b, 3]) == np.sign(G[1][b][i]))
            if hit:
                ok.append(j)
        assert ok, f"query {b}: result {par[b].tolist()} is no draw of its own query for any counter"
        cands.append(ok)
        if k == 1:
            assert par[b, 4] == run["fixed"][1][b, 4]
        if k == 3:
            assert abs(par[b, 3]) == abs(run["fixed"][3][b, 3]) and par[b, 4] == run["fixed"][3][b, 4]
    assert _perfect_matching(cands) == B, "no one-to-one assignment of counters to queries reproduces the results"
