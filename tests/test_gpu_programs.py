"""Programs written against the drop-in boundary, run on the device.

* bin/itebd_dropin: an iTEBD program written only against the C++ API in the
  call pattern of the reference's example (iTEBD.cpp:148-216), compiled
  against include/uni10_amd/uni10.hpp; its energy and Schmidt spectra after
  the run match the oracle's iTEBD to 1e-10.
* bin/seam_check: the backend seam with the reference's own signatures
  (include/uni10_seam_amd.h) called the way the reference's classes call it,
  with device and host operands; every result matches numpy on the same
  seeded inputs (bit-exact for copies, permutes and diagonal scalings).
"""
import os
import subprocess

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "trim_amd", "bin")
NETS = os.path.join(ROOT, "trim_amd", "nets")


def _run(args, timeout=120):
    print("run %s" % " ".join(os.path.basename(a) for a in args[:1]) + " " + " ".join(args[1:]), flush=True)
    r = subprocess.run(args, stdout=subprocess.PIPE, stderr=subprocess.PIPE, text=True, timeout=timeout)
    assert r.returncode == 0, r.stderr
    return r.stdout


@pytest.mark.parametrize("chi,steps,delta", [(8, 30, 1.0), (32, 20, 0.6)])
def test_itebd_program_against_the_api_matches_oracle(ora, golden, chi, steps, delta):
    out = _run([os.path.join(BIN, "itebd_dropin"), str(chi), str(steps), "7", "0.01", repr(delta), NETS])
    vals = {l.split()[0]: [float(v) for v in l.split()[1:]] for l in out.splitlines() if l.strip()}
    O = ora.ITEBD(chi, 2, ora.heisenberg_two_site(delta), 0.01, 7, golden)
    e0 = O.energy()
    for _ in range(steps):
        O.step()
    assert abs(vals["energy0"][0] - e0) <= 1e-12
    assert abs(vals["energy"][0] - O.energy()) <= 1e-10
    assert np.max(np.abs(np.array(vals["lambdaA"]) - O.La)) <= 1e-10
    assert np.max(np.abs(np.array(vals["lambdaB"]) - O.Lb)) <= 1e-10


def test_seam_functions_match_numpy(ora, tmp_path):
    _run([os.path.join(BIN, "seam_check"), str(tmp_path)])
    r = {f[:-4]: np.fromfile(str(tmp_path / f), dtype=np.float64) for f in os.listdir(tmp_path)}
    u = ora.splitmix64_uniform
    M, K, N = 37, 29, 23
    a, b = u(1, M * K).reshape(M, K), u(2, K * N).reshape(K, N)
    ab = a @ b
    for k in ("mm_dev", "mm_host"):
        assert np.linalg.norm(r[k].reshape(M, N) - ab) <= 1e-14 * np.linalg.norm(ab)
    s = u(3, 40 * 31).reshape(40, 31)
    sv = np.linalg.svd(s, compute_uv=False)
    U, S, VT = r["svd_U"].reshape(40, 31), r["svd_S"], r["svd_VT"].reshape(31, 31)
    assert np.allclose(S, sv, rtol=0, atol=1e-13 * sv[0]) and np.allclose(r["svd_S_host"], sv, rtol=0,
                                                                             atol=1e-13 * sv[0])
    assert np.allclose(U @ np.diag(S) @ VT, s, rtol=0, atol=1e-13)
    x = u(4, 144).reshape(12, 12)
    h = x + x.T
    w, v = r["eig_w"], r["eig_v"].reshape(12, 12)
    assert np.allclose(w, np.linalg.eigvalsh(h), rtol=0, atol=1e-13)
    assert np.allclose(v @ h @ v.T, np.diag(w), rtol=0, atol=1e-12)          # eigenvectors are rows
    y0, x0 = u(5, 100), u(6, 100)
    assert np.allclose(r["norm"], [np.linalg.norm(x0), np.linalg.norm(x0[:100:3][:34]), np.linalg.norm(x0[::2])],
                       rtol=1e-15, atol=0)
    assert np.allclose(r["sum"], [x0.sum(), x0[::2].sum(), x0.sum()], rtol=0, atol=1e-13)
    assert r["add"].tobytes() == (y0 + x0).tobytes()
    assert r["scal"].tobytes() == (-2.5 * (y0 + x0)).tobytes()
    assert r["mul"].tobytes() == (y0 * x0).tobytes()
    dr, dc = u(7, M), u(8, K)
    assert r["diagrow"].tobytes() == (a * dr[:, None]).reshape(-1).tobytes()
    assert r["diagcol"].tobytes() == (a * dc[None, :]).reshape(-1).tobytes()
    assert r["transpose"].tobytes() == np.ascontiguousarray(a.T).tobytes()
    assert r["transpose_inplace"].tobytes() == np.ascontiguousarray(a.T).tobytes()
    assert r["identity"].tobytes() == np.eye(5, 7).reshape(-1).tobytes()
    g, dg = u(9, 35).reshape(5, 7).copy(), u(10, 5)
    g2 = g.copy()
    g2[np.arange(5), np.arange(5)] = dg
    assert r["setdiag"].tobytes() == g2.reshape(-1).tobytes()
    assert r["getdiag"].tobytes() == np.diag(g)[:4].tobytes()
    t = u(11, 360).reshape(5, 3, 4, 6)
    assert r["reshape"][:7].tobytes() == np.transpose(t, (2, 0, 3, 1)).reshape(-1)[:7].tobytes()
    assert r["reshape"][9:].tobytes() == np.transpose(t, (2, 0, 3, 1)).reshape(-1)[9:].tobytes()
    assert r["elemat"][0] == 42.0 and r["elemat"][1] == np.transpose(t, (2, 0, 3, 1)).reshape(-1)[8]
    xl = u(12, 3600).reshape(60, 60)
    al = xl + xl.T
    e0, iters, conv = r["lanczos"]
    assert conv == 1.0 and abs(e0 - np.linalg.eigvalsh(al)[0]) <= 1e-12 * np.abs(np.linalg.eigvalsh(al)).max()
    vec = r["lanczos_vec"]
    assert np.linalg.norm(al @ vec - e0 * vec) <= 1e-6 * np.linalg.norm(vec)


def test_engines_default_to_the_packaged_network_files(golden):
    """The iTEBD drivers read their .net files from the package (beside
    libuni10_amd.so) when no directory is given: same numbers as with the
    reference's files given explicitly."""
    from trim_amd import uni10
    uni10.load()
    a = uni10.ITebd(8, 2, 1.0, 0.01, 11)
    b = uni10.ITebd(8, 2, 1.0, 0.01, 11, golden)
    a.step(3)
    b.step(3)
    assert a.energy() == b.energy()
    B = uni10.BatchTebd([8, 16], [1.0, 0.5], [3, 4], 0.01)
    B.step(2)
    assert np.isfinite(B.energy(1))
    for f in ("itebd.net", "updateA.net", "MPS.net", "measure.net"):
        assert open(os.path.join(NETS, f), "rb").read() == open(os.path.join(golden, f), "rb").read()
