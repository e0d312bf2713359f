"""CPU-side checks of the native boundary (no GPU calls).

* both libraries build and load;
* every function declared in include/tn_hip.h and include/uni10_capi.h is
  exported by the matching library and has a ctypes signature in the binding;
* the host C++ layer links only against the C ABI (libtn_hip.so), not HIP
  headers of the kernel library internals;
* calls that need the device fail with an error code (never a CPU fallback)
  when no GPU is present.
"""
import ctypes
import os
import re
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _decls(header):
    text = open(os.path.join(ROOT, "include", header)).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b((?:tn|u10)_[a-z0-9_]+)\s*\(", text)))


@pytest.fixture(scope="module")
def built():
    from trim_amd import _build
    if not (os.path.exists(_build.TN_LIB) and os.path.exists(_build.U10_LIB)):
        _build.build(verbose=False)
    return _build


def _exports(path):
    out = subprocess.run(["nm", "-D", "--defined-only", path], stdout=subprocess.PIPE, text=True).stdout
    return set(l.split()[-1] for l in out.splitlines() if " T " in l)


def test_tn_hip_exports_every_declared_symbol(built):
    decl = _decls("tn_hip.h")
    assert len(decl) >= 40
    exp = _exports(built.TN_LIB)
    missing = [d for d in decl if d not in exp]
    assert not missing, missing
    from trim_amd import tn
    assert sorted(tn.SIGNATURES) == decl


def test_uni10_capi_exports_every_declared_symbol(built):
    decl = _decls("uni10_capi.h")
    exp = _exports(built.U10_LIB)
    missing = [d for d in decl if d not in exp]
    assert not missing, missing
    from trim_amd import uni10
    assert sorted(uni10.SIGNATURES) == decl


def test_host_layer_depends_only_on_the_c_abi(built):
    deps = subprocess.run(["ldd", built.U10_LIB], stdout=subprocess.PIPE, text=True).stdout
    assert "libtn_hip.so" in deps
    src_dir = os.path.join(ROOT, "trim_amd", "csrc", "uni10")
    for f in os.listdir(src_dir):
        text = open(os.path.join(src_dir, f)).read()
        assert "hip_runtime" not in text and "tn_internal.h" not in text, f


def test_libraries_load_and_fail_loudly_without_device(built):
    from trim_amd import tn
    L = tn.load()
    import torch  # noqa: F401  (the image's torch sees no device here)
    if torch.cuda.is_available():
        pytest.skip("a GPU is present")
    rc = L.tn_init(0)
    assert rc != 0
    assert L.tn_last_error().decode()   # a message, not a silent fallback
    p = ctypes.c_void_p()
    assert L.tn_malloc(ctypes.byref(p), 64) != 0


def test_no_oracle_in_product():
    """The product never imports or links the oracle."""
    for dirpath, _, files in os.walk(os.path.join(ROOT, "trim_amd")):
        for f in files:
            if f.endswith((".py", ".cpp", ".hip", ".h")):
                text = open(os.path.join(dirpath, f)).read()
                assert "oracle" not in text.lower().replace("oracle-free", ""), os.path.join(dirpath, f)


CALL_SITES = r'''
// The reference's own call sites for the hot path, unchanged
// (iTEBD.cpp:148-216, UniTensor.h:583-592,973, Network.h:74-127, Block.h:389)
#include "uni10_amd/uni10.hpp"
using namespace uni10;
int main() {
  std::vector<Bond> bonds = {Bond(BD_IN, 4), Bond(BD_IN, 2), Bond(BD_OUT, 4)};
  UniTensor A(bonds, "A"), B(bonds, "B");
  std::vector<int> la = {0, 1, 2}, lb = {2, 3, 4};
  A.setLabel(la);
  B.setLabel(lb);
  UniTensor C = contract(A, B, true);
  std::vector<int> order = {4, 1, 0, 3};
  C.permute(order, 2);
  Network net("itebd.net");
  net.putTensor("ALa", &A);
  net.putTensor("BLb", B);
  UniTensor T = net.launch();
  Matrix M = T.getBlock();
  std::vector<Matrix> usv = M.svd();
  std::vector<Matrix> eig = usv[0].eigh();
  Matrix psi(M.row(), 1);
  double E0 = 0.0;
  size_t it = lanczosEigh(eig[1], E0, psi, 100, 1e-12);
  return (int)it + (int)T.elemNum() + (int)takeExp(-0.01, eig[1]).row();
}
'''


def test_cpp_api_compiles_and_links_as_the_reference_call_sites(built, tmp_path):
    # the C++ drop-in keeps the reference's signatures: a program written
    # against them builds and links against libuni10_amd.so (not run here:
    # no GPU in this container)
    src = tmp_path / "call_sites.cpp"
    src.write_text(CALL_SITES)
    exe = tmp_path / "call_sites"
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    pkg = os.path.dirname(built.U10_LIB)
    r = subprocess.run(["g++", "-std=c++17", "-I", os.path.join(root, "include"), str(src), "-o", str(exe),
                        "-L", pkg, "-luni10_amd", "-ltn_hip", "-Wl,-rpath," + pkg],
                       stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    assert r.returncode == 0, r.stdout


def test_product_bond_combine_known_answer_and_oracle(built, ora):
    """The product's Bond::combine (host code, no device) against
    testBond.cpp:101-160 (dimension 24, degeneracies -2:1, -1:5, 0:9, 1:7,
    2:2) and, for the free combine(tp, bonds) over mixed in/out bonds,
    against the oracle's run-length result and degeneracy map."""
    from trim_amd import uni10 as u
    u.load()
    bd = u.Bond(u.BD_IN, [(1, 0, 0), (1, 0, 0), (0, 0, 0), (0, 0, 0), (0, 0, 0), (-1, 0, 0)])
    bd2 = u.Bond(u.BD_IN, [(1, 0, 0), (0, 0, 0), (0, 0, 0), (-1, 0, 0)])
    tp, dim, runs, deg = u.combine_bonds([bd2, bd])
    assert tp == u.BD_IN and dim == 24
    assert [(q.U1, n) for q, n in deg] == [(-2, 1), (-1, 5), (0, 9), (1, 7), (2, 2)]
    Q = ora.Qnum
    spec = [(ora.BD_IN, [Q(1), Q(-1)]), (ora.BD_OUT, [Q(2), Q(0), Q(0), Q(1, 0, 1)]),
            (ora.BD_IN, [Q(1), Q(0), Q(-1), Q(0, 1, 0)])]
    obonds = [ora.Bond(t, qs) for t, qs in spec]
    pbonds = [u.Bond(t, [q.triple() for q in qs]) for t, qs in spec]
    for ttp in (ora.BD_IN, ora.BD_OUT):
        oc = ora.combine_bonds(ttp, obonds)
        tp, dim, runs, deg = u.combine_bonds(pbonds, tp=ttp)
        assert tp == oc.type and dim == oc.dim
        assert [(q.triple(), n) for q, n in runs] == [(q.triple(), n) for q, n in zip(oc.Qnums, oc.Qdegs)]
        assert [(q.triple(), n) for q, n in deg] == [(q.triple(), n) for q, n in oc.degeneracy().items()]


def test_seam_library_exports_the_reference_signatures(built):
    """libuni10_seam_amd.so defines every function of include/uni10_seam_amd.h
    with the reference's C++ signature (matched on the demangled names), and
    links only the C ABI library."""
    hdr = open(os.path.join(ROOT, "include", "uni10_seam_amd.h")).read()
    hdr = re.sub(r"/\*.*?\*/", "", hdr, flags=re.S)
    hdr = re.sub(r"//[^\n]*", "", hdr)
    decls = re.findall(r"\b([A-Za-z_][A-Za-z0-9_]*)\(([^)]*)\);", hdr)
    assert len(decls) >= 24
    lib = os.path.join(os.path.dirname(built.TN_LIB), "libuni10_seam_amd.so")
    out = subprocess.run(["nm", "-D", "-C", "--defined-only", lib], stdout=subprocess.PIPE, text=True).stdout
    exported = set(l.split(" T ", 1)[1] for l in out.splitlines() if " T " in l)
    for name, args in decls:
        types = []
        for a in args.split(","):
            a = a.strip()
            t = re.sub(r"\s*\b[A-Za-z_][A-Za-z0-9_]*$", "", a)       # drop the parameter name
            t = t.replace(" *", "*").replace(" &", "&")
            types.append(re.sub(r"^const (\w+)\*$", r"\1 const*", t))    # the demangler's spelling
        sig = "uni10::%s(%s)" % (name, ", ".join(types))
        sig = sig.replace("size_t", "unsigned long")
        assert sig in exported, sig
    deps = subprocess.run(["ldd", lib], stdout=subprocess.PIPE, text=True).stdout
    assert "libtn_hip.so" in deps and "libuni10_amd.so" not in deps


def test_descriptor_layouts_match_the_c_header(tmp_path):
    """The ctypes binding's numpy records for the GEMM descriptors and index
    maps have the size and field offsets of the C structs in tn_hip.h (a C
    program built with gcc against the header prints them)."""
    from trim_amd import tn
    fields = {"tn_gemm_desc": tn.GEMM_DESC, "tn_index_map": tn.INDEX_MAP, "tn_gemm_desc_ex": tn.GEMM_DESC_EX}
    src = ["#include <stdio.h>", "#include <stddef.h>", '#include "tn_hip.h"', "int main(void) {"]
    for st, dt in fields.items():
        src.append('printf("%s %%zu\\n", sizeof(%s));' % (st, st))
        for name in dt.names:
            src.append('printf("%s.%s %%zu\\n", offsetof(%s, %s));' % (st, name, st, name))
    src.append("return 0; }")
    c = tmp_path / "layout.c"
    c.write_text("\n".join(src))
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-std=c11", "-I", os.path.join(ROOT, "include"), str(c), "-o", str(exe)], check=True)
    out = dict(l.split() for l in subprocess.run([str(exe)], stdout=subprocess.PIPE, text=True,
                                                 check=True).stdout.splitlines())
    for st, dt in fields.items():
        assert int(out[st]) == dt.itemsize, st
        for name in dt.names:
            assert int(out["%s.%s" % (st, name)]) == dt.fields[name][1], (st, name)
