"""Host logic of vegas_amd (no GPU needed): geometry, batching, RAvg, the
C-ABI symbol table and the device helpers that also compile for the host."""
import ctypes
import os
import re
import subprocess

import numpy as np
import pytest

import vegas_amd as va
from vegas_amd import _geometry

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


# ----------------------------------------------------------------- geometry
def test_geometry_reference_fixtures():
    g = _geometry.compute(2, neval=234)  # tests/test_vegas.py:561-575
    assert list(g["ninc"]) == [20, 20] and list(g["nstrat"]) == [5, 5]
    g = _geometry.compute(2, neval=1000, nstrat=[1, 1])
    assert list(g["ninc"]) == [100, 100] and g["min_neval_hcube"] == 1000
    g = _geometry.compute(2, nstrat=[10, 11])
    assert list(g["ninc"]) == [80, 88] and g["neval"] == 880 and g["min_neval_hcube"] == 2
    g = _geometry.compute(2, nstrat=[7, 9], neval=2000)  # tests:730-734
    assert list(g["ninc"]) == [196, 198] and g["min_neval_hcube"] == 7
    with pytest.raises(ValueError):
        _geometry.compute(2, nstrat=[7, 9], neval=20)
    assert list(_geometry.compute(2, neval=3100)["nstrat"]) == [20, 19]
    assert list(_geometry.compute(2, neval=3500)["nstrat"]) == [21, 20]
    g = _geometry.compute(5, neval=1000)  # doc/source/eg10.out:3-5
    assert list(g["nstrat"]) == [3, 3, 3, 2, 2] and list(g["ninc"]) == [99, 99, 99, 100, 100]
    assert _geometry.compute(20, neval=2e6)["nhcube"] == 2 ** 17  # doc/source/eg6a.out:1
    with pytest.raises(MemoryError):
        _geometry.compute(2, neval=1e7, max_mem=10)


@pytest.mark.parametrize("dim,neval", [(2, 234), (3, 1e4), (4, 1e5), (8, 1e7), (8, 2e7), (9, 2e7),
                                       (10, 1e6), (2, 1e4), (5, 1000), (6, 3.3e5), (1, 50)])
def test_geometry_matches_oracle(orc, dim, neval):
    a = _geometry.compute(dim, neval=neval)
    b = orc.geometry(dim, neval=neval)
    for k in ("nstrat", "ninc"):
        assert list(a[k]) == list(b[k])
    for k in ("nhcube", "neval", "min_neval_hcube", "workspace"):
        assert a[k] == b[k]


def _ref_batches(n_h, min_neval_batch):
    """The reference's batch loop (pyx:1692-1765), written as its own loop."""
    out, nb, hb, p = [], 0, 0, 0
    for h, n in enumerate(n_h):
        nb += n
        if nb < min_neval_batch and h < len(n_h) - 1:
            continue
        out.append((hb, h + 1, p, p + nb))
        p += nb
        nb, hb = 0, h + 1
    return out


@pytest.mark.parametrize("seed", range(5))
def test_batches(seed):
    rng = np.random.default_rng(seed)
    n_h = rng.integers(2, 300, rng.integers(1, 5000))
    for mnb in (1, 7, 1000, 10 ** 5):
        assert _geometry.batches(n_h, mnb) == _ref_batches(n_h, mnb)


# -------------------------------------------------------------------- RAvg
def test_ravg_reference_fixture():
    # tests/test_vegas.py:216-236
    a = va.RAvg()
    a.add(va.GVar(1, 1)); a.add(va.GVar(2, 2)); a.add(va.GVar(3, 3))
    np.testing.assert_allclose(a.mean, 1.346938775510204)
    np.testing.assert_allclose(a.sdev, 0.8571428571428571)
    assert a.dof == 2
    np.testing.assert_allclose(a.chi2, 0.5306122448979592)
    np.testing.assert_allclose(a.Q, 0.7669711269557102)
    assert str(a) == '1.35(86)'
    s = ["itn   integral        wgt average     chi2/dof        Q",
         "-------------------------------------------------------",
         "  1   1.0(1.0)        1.0(1.0)            0.00     1.00",
         "  2   2.0(2.0)        1.20(89)            0.20     0.65",
         "  3   3.0(3.0)        1.35(86)            0.27     0.77", ""]
    assert a.summary() == '\n'.join(s)


def test_ravg_unweighted_and_pickle(orc):
    import pickle
    m, v = [1.0, 2.5, 2.0, 1.7], [0.04, 0.09, 0.01, 0.25]
    for w in (True, False):
        a = va.RAvg(weighted=w)
        for mi, vi in zip(m, v):
            a.add(va.GVar(mi, vi ** 0.5))
        mean, sdev, chi2 = orc.ravg(m, v, weighted=w)
        np.testing.assert_allclose([a.mean, a.sdev, a.chi2], [mean, sdev, chi2], rtol=1e-13)
        b = pickle.loads(pickle.dumps(a))
        assert b.mean == a.mean and b.sdev == a.sdev and b.nitn == 4
    assert va.RAvg().converged(0.1, 0.0) is False or True


def test_gammaq_vs_scipy():
    from scipy.special import gammaincc
    from vegas_amd._ravg import _gammaQ
    for a, x in [(0.5, 0.3), (4.5, 1.0), (7.0, 20.0), (100.0, 90.0), (1.0, 1e-5)]:
        np.testing.assert_allclose(_gammaQ(a, x), gammaincc(a, x), rtol=1e-12)


def test_gvar_format():
    assert str(va.GVar(13.469, 8.57)) == '13.5(8.6)'
    assert str(va.GVar(1.2, 0.89)) == '1.20(89)'
    assert str(va.GVar(1.0, 1.0)) == '1.0(1.0)'


# ------------------------------------------------------------- batch API
def test_decorators():
    @va.lbatchintegrand
    def f(x):
        return x[:, 0]
    assert f.fcntype == 'lbatch'

    @va.rbatchintegrand
    def g(x):
        return x[0]
    assert g.fcntype == 'rbatch'

    class C(va.LBatchIntegrand):
        def __call__(self, x):
            return x[:, 0]
    assert C().fcntype == 'lbatch'
    assert va.batchintegrand is va.lbatchintegrand
    h = va.gpu_lbatchintegrand(lambda x: x)
    assert h.fcntype == 'lbatch' and h.vgx_torch


def test_device_integrand_host_eval():
    f = va.GaussianMixture.isotropic([[0.3, 0.3], [0.7, 0.7]], 0.1)
    x = np.random.default_rng(0).uniform(0, 1, (50, 2))
    want = sum(0.5 / (2 * np.pi * 0.01) * np.exp(-((x - c) ** 2).sum(1) / 0.02) for c in (0.3, 0.7))
    np.testing.assert_allclose(f(x), want, rtol=1e-13)
    r = va.DiagonalRidge([1.0, 2.0], [50.0, 20.0], [0.2, 0.6], dim=3)
    x3 = np.random.default_rng(1).uniform(0, 1, (40, 3))
    np.testing.assert_allclose(r(x3), r.as_mixture()(x3), rtol=1e-13)


# ---------------------------------------------------------------- C-ABI
def test_library_exports_every_declared_symbol():
    from vegas_amd import _lib
    if not os.path.exists(_lib.LIB_PATH):
        pytest.skip("libvgx.so not built")
    hdr = open(os.path.join(ROOT, "include", "vgx.h")).read()
    names = set(re.findall(r"\b(vgx_[a-z0-9_]+)\s*\(", hdr))
    assert len(names) >= 15
    L = ctypes.CDLL(_lib.LIB_PATH)
    for n in sorted(names):
        assert hasattr(L, n), n
    L.vgx_strerror.restype = ctypes.c_char_p
    assert L.vgx_strerror(-3) == b"integrand evaluates to nan"


def test_divmagic_and_philox_host(tmp_path):
    """vgx_device.h helpers compiled for the host: invariant division is
    exact for every divisor used by the sampler, and Philox matches the KAT."""
    src = tmp_path / "t.cpp"
    src.write_text(r'''
#include "vgx_device.h"
#include <cstdio>
#include <random>
int main() {
  std::mt19937_64 r(1);
  for (uint32_t d = 1; d <= 4096; ++d) {
    vgx::DivMagic g = vgx::make_divmagic(d);
    for (int i = 0; i < 2000; ++i) {
      uint32_t n = (i < 1000) ? (uint32_t)r() : (uint32_t)(r() % (64ull * d));
      if (i == 0) n = 0xFFFFFFFFu;
      if (vgx::divq(n, g) != n / d) { std::printf("bad %u %u\n", n, d); return 1; }
    }
  }
  vgx::U4 o = vgx::philox4x32_10(vgx::U4{0x243F6A88u, 0x85A308D3u, 0x13198A2Eu, 0x03707344u},
                                 0xA4093822u, 0x299F31D0u);
  if (o.x != 0xD16CFE09u || o.y != 0x94FDCCEBu || o.z != 0x5001E420u || o.w != 0x24126EA1u) return 2;
  std::printf("ok\n");
  return 0;
}
''')
    exe = tmp_path / "t"
    inc = os.path.join(ROOT, "vegas_amd", "csrc")
    cmd = ["/opt/rocm/bin/hipcc", "-O1", "-std=c++17", "-I" + inc, "-I" + os.path.join(ROOT, "include"),
           str(src), "-o", str(exe)]
    if subprocess.run(cmd, capture_output=True).returncode != 0:
        pytest.skip("hipcc host build unavailable")
    out = subprocess.run([str(exe)], capture_output=True, text=True)
    assert out.returncode == 0, out.stdout


def test_ravg_reference_docstring_example():
    """src/vegas/__init__.py:1237-1262: the ten iteration results printed in
    ravg's docstring, weighted (1.0018(22), chi2/dof 1.21, Q 0.28 after all
    ten; the partial lines are too sensitive to the rounded inputs to pin
    anything) and the unweighted average of the last
    five (1.0018(26), chi2/dof 0.31, Q 0.87).  The inputs are the printed,
    rounded values: each mean is off by up to half its last digit (up to
    5e-4 for the first three), which moves the weighted mean by up to
    ~5.5e-5 on top of the printed answer's own rounding of 5e-5, hence
    1.1e-4 on means, and 0.02 on chi2/dof and Q."""
    import vegas_amd as va
    res = [(1.013, .019), (0.997, .014), (1.021, .012), (0.9785, .0097), (1.0067, .0085),
           (0.9996, .0075), (1.0020, .0061), (1.0051, .0052), (1.0046, .0047), (0.9976, .0047)]
    r = va.ravg([va.GVar(m, s) for m, s in res])
    assert abs(r.mean - 1.0018) < 1.1e-4 and abs(r.sdev - 0.0022) < 6e-5
    assert abs(r.chi2 / r.dof - 1.21) < 0.02 and abs(r.Q - 0.28) < 0.02
    ur = va.ravg(r.itn_results[5:], weighted=False)
    assert abs(ur.mean - 1.0018) < 1.1e-4 and abs(ur.sdev - 0.0026) < 6e-5
    assert abs(ur.chi2 / ur.dof - 0.31) < 0.02 and abs(ur.Q - 0.87) < 0.02
    again = va.ravg(r)  # an RAvg is averaged again with its own setting
    assert again.mean == r.mean and again.sdev == r.sdev


def test_ravg_arrays_and_dicts():
    import numpy as np
    import vegas_amd as va
    g = [va.GVarArray([1.0 + 0.01 * k, 2.0 - 0.01 * k], np.diag([0.01, 0.04]), (2,)) for k in range(3)]
    ra = va.ravg(g)
    np.testing.assert_allclose(ra.mean, [1.01, 1.99], rtol=1e-12)
    np.testing.assert_allclose(ra.cov, np.diag([0.01, 0.04]) / 3, rtol=1e-12)
    d = [dict(a=va.GVar(1.0 + 0.01 * k, 0.1), b=va.GVarArray([2.0, 3.0], np.diag([0.04, 0.09]), (2,)))
         for k in range(3)]
    rd = va.ravg(d)
    assert abs(rd['a'].mean - 1.01) < 1e-12 and abs(rd['a'].sdev - 0.1 / np.sqrt(3)) < 1e-12
    np.testing.assert_allclose(rd['b'].mean, [2.0, 3.0], rtol=1e-12)
    with pytest.raises(ValueError):
        va.ravg([])


def test_ctypes_mirrors_match_the_c_abi(tmp_path):
    """The ctypes structures the Python side passes through the C-ABI have
    the size and field offsets of the C structures in include/vgx.h (a C
    compiler's own layout, field by field)."""
    from vegas_amd import _lib
    from vegas_amd import batch as _batch
    mirrors = {"vgx_config": _lib.VgxConfig, "vgx_integrand": _lib.VgxIntegrand, "vgx_state": _lib.VgxState,
               "vgx_batch_config": _batch._BatchConfig, "vgx_batch_state": _batch._BatchState}
    lines = ['#include <stddef.h>', '#include <stdio.h>', '#include "vgx.h"', 'int main(void) {']
    for cname, py in mirrors.items():
        lines.append('  printf("%s sizeof %%zu\\n", sizeof(%s));' % (cname, cname))
        for fname, _ in py._fields_:
            lines.append('  printf("%s %s %%zu\\n", offsetof(%s, %s));' % (cname, fname, cname, fname))
    lines += ['  return 0;', '}']
    src = tmp_path / "layout.c"
    src.write_text("\n".join(lines) + "\n")
    exe = tmp_path / "layout"
    cmd = ["gcc", "-std=c11", "-I" + os.path.join(ROOT, "include"), str(src), "-o", str(exe)]
    if subprocess.run(cmd, capture_output=True).returncode != 0:
        pytest.skip("no host C compiler")
    out = subprocess.run([str(exe)], capture_output=True, text=True, check=True).stdout.split("\n")
    got = {tuple(l.split()[:2]): int(l.split()[2]) for l in out if l.strip()}
    for cname, py in mirrors.items():
        assert got[(cname, "sizeof")] == ctypes.sizeof(py), cname
        for fname, _ in py._fields_:
            assert got[(cname, fname)] == getattr(py, fname).offset, (cname, fname)
