"""The C-ABI library and its host-side pieces, without a GPU."""
import re
from pathlib import Path

import numpy as np
import pytest

from gms_liberi_amd import census, configs, liberi
from gms_liberi_amd.caseio import Case, pack, unpack

ROOT = Path(__file__).resolve().parents[1]
HEADER = ROOT / "include" / "liberi_hip.h"


def declared_functions():
    text = re.sub(r"/\*.*?\*/", "", HEADER.read_text(), flags=re.S)
    return sorted(set(re.findall(r"\b(liberi_c_\w+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    lib = liberi.load_library()
    names = declared_functions()
    assert len(names) >= 14
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def test_python_mirror_matches_the_header():
    """Option numbers and the liberi_stats layout of the Python mirror are the
    header's (a field missing in the mirror would shift every later one)."""
    text = re.sub(r"/\*.*?\*/", "", HEADER.read_text(), flags=re.S)
    opts = dict(re.findall(r"#define LIBERI_(OPT_\w+)\s+(\d+)", text))
    assert len(opts) >= 14
    for name, val in opts.items():
        assert getattr(liberi, name) == int(val), name
    body = re.search(r"typedef struct\s*\{(.*?)\}\s*liberi_stats;", text, flags=re.S).group(1)
    fields = re.findall(r"\b(\w+)(?:\[\w+\])?;", body)
    assert [f[0] for f in liberi.Stats._fields_] == fields


def test_class_tables_agree(oracle):
    assert liberi.class_names() == census.NAMES == oracle.class_names
    assert len(census.CLASS_L) == 55


def test_create_without_device_reports_edevice():
    try:
        import torch

        if torch.cuda.is_available():
            pytest.skip("a GPU is visible")
    except ImportError:
        pass
    with pytest.raises(liberi.LiberiError) as e:
        liberi.Handle()
    assert e.value.code == 4
    assert "device" in str(e.value)


def test_last_error_is_set_by_null_arguments():
    lib = liberi.load_library()
    assert lib.liberi_c_create(None) == 1
    assert "NULL" in lib.liberi_c_last_error().decode()
    assert lib.liberi_c_fock_build(None, None, None, 0) == 1
    assert lib.liberi_c_fock(None, None, 0) == 2  # before liberi_init
    assert "liberi_init" in lib.liberi_c_last_error().decode()


def test_census_counts_are_consistent():
    for c in range(55):
        p = census.class_params(c)
        assert p["nr"] == (p["la"] + p["lb"] + p["lc"] + p["ld"]) // 2 + 1
        assert census.flops_per_prim_quartet(c) > 0 and census.flops_per_quartet(c) > 0


def test_pack_unpack_roundtrip():
    rng = np.random.default_rng(0)
    m = rng.standard_normal((7, 7))
    m = m + m.T
    p = pack(m)
    assert p.shape == (28,)
    assert p[2] == m[1, 1] and p[1] == m[1, 0]
    assert np.array_equal(unpack(p), m)


def test_case_stream_roundtrip(tmp_path):
    c = Case.load_npz(ROOT / "tests" / "golden" / "cfg_water.npz")
    c.write_stream(tmp_path / "w.case")
    d = Case.read_stream(tmp_path / "w.case", "w")
    assert np.array_equal(d.basis.schwrz, c.basis.schwrz)
    assert np.array_equal(d.density, c.density)
    assert np.array_equal(d.basis.sh_loc, c.basis.sh_loc)


@pytest.mark.parametrize("name,nbf,nsh", [("water", 13, 9), ("ethane", 42, 24), ("c60", 900, 360),
                                          ("cn_chain", 620, 160)])
def test_workload_generators(name, nbf, nsh):
    c = configs.GENERATORS[name]()
    assert (c.basis.num_bas, c.basis.nsh) == (nbf, nsh)
    # density = 2 C C^T with orthonormal C: (D/2)^2 = D/2, trace = 2 n_occ
    D = unpack(c.density)
    assert np.allclose((D / 2) @ (D / 2), D / 2, atol=1e-10)
    # committed fixture holds exactly these inputs
    g = Case.load_npz(ROOT / "tests" / "golden" / f"cfg_{name}.npz")
    assert np.array_equal(g.basis.exponents, c.basis.exponents)
    assert np.array_equal(g.basis.coef_s, c.basis.coef_s)
    assert np.allclose(g.density, c.density, atol=1e-12)


def test_c60_geometry():
    p = configs.truncated_icosahedron(3.55)
    r = np.linalg.norm(p, axis=1)
    assert np.allclose(r, 3.55)
    d = np.linalg.norm(p[:, None] - p[None], axis=2) + np.eye(60) * 99
    nn = np.sort(d, axis=1)[:, :3]  # three neighbours per vertex
    assert np.all(nn < 1.5) and np.all(nn > 1.3)


def test_contraction_normalisation():
    # contracted functions have unit self-overlap in the component the GAMESS
    # factors normalise (x for p, xy for d, xyz for f)
    c = configs.cn_chain()
    b = c.basis
    for s in range(b.nsh):
        l = b.ang_mom[s] - 1
        k = b.contr_num[s]
        o = b.sh_loc[s] - 1
        e = b.exponents[o:o + k]
        w = [b.coef_s, b.coef_p, b.coef_d, b.coef_f][l][o:o + k]
        p = e[:, None] + e[None]
        ov = (np.outer(w, w) * (np.pi / p) ** 1.5 * (0.5 / p) ** l).sum()
        assert abs(ov - 1.0) < 1e-12


def test_group_chunk_order_is_a_permutation(tmp_path):
    """Workgroup b of a group-kernel launch of `grid` workgroups takes chunk
    b * chunk_stride(grid) mod grid: a bijection for every grid the launcher
    can use (1 .. kGroupWorkgroups), including grids equal to or sharing
    factors with any fixed multiplier."""
    import shutil
    import subprocess

    gxx = shutil.which("g++")
    if gxx is None:
        pytest.skip("g++ not available")
    src = tmp_path / "chunk.cpp"
    src.write_text(r'''
#include <cstdio>
#include <vector>
#include "liberi_internal.hpp"
int main() {
  std::vector<unsigned> seen(liberi::kGroupWorkgroups + 1, 0);
  for (int64_t grid = 1; grid <= liberi::kGroupWorkgroups; ++grid) {
    const int64_t s = liberi::chunk_stride(grid);
    if (s < 1 || (grid > 1 && s >= grid)) { std::printf("bad stride %lld for %lld\n", (long long)s, (long long)grid); return 1; }
    for (int64_t b = 0; b < grid; ++b) {
      const int64_t c = b * s % grid;
      if (seen[c] == (unsigned)grid) { std::printf("grid %lld: chunk %lld taken twice\n", (long long)grid, (long long)c); return 1; }
      seen[c] = (unsigned)grid;
    }
  }
  std::printf("ok\n");
  return 0;
}
''')
    exe = tmp_path / "chunk"
    subprocess.run([gxx, "-O2", "-std=c++17", "-I", str(ROOT / "gms_liberi_amd" / "csrc"), str(src), "-o", str(exe)],
                   check=True)
    r = subprocess.run([str(exe)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0 and r.stdout.strip() == "ok", r.stdout
