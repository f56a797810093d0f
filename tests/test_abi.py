"""The C-ABI library loads without a GPU and exports every entry point that
include/xavs2_amd.h declares; the ctypes mirror binds the same names.  No
compute call is made here."""
import ctypes as C
import os
import re
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "xavs2_amd.h")
LIB = os.path.join(ROOT, "xavs2_amd", "libxavs2_amd.so")


def declared():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", " ", text, flags=re.S)
    names = re.findall(r"\b(xme_[a-z0-9_]+)\s*\(", text)
    # typedef'd function pointer types are not entry points
    types = set(re.findall(r"\(\s*\*\s*(xme_[a-z0-9_]+)\s*\)", text))
    return sorted(set(names) - types)


def exported():
    out = subprocess.run(["nm", "-D", "--defined-only", LIB], capture_output=True, text=True, check=True).stdout
    return {line.split()[-1] for line in out.splitlines() if " T " in line}


def test_header_declares_the_boundary():
    names = declared()
    # the four g_funcs tables, the batched primitives, the frame job, the
    # reference-plane producer, MC, the residual path and the server
    for must in ("xme_get_sad_table", "xme_get_satd_table", "xme_get_sad_x3_table", "xme_get_sad_x4_table",
                 "xme_sad_batch", "xme_satd_batch", "xme_sad_x3_batch", "xme_sad_x4_batch", "xme_plan_create",
                 "xme_me_run", "xme_me_run_graph", "xme_interpolate_frame", "xme_expand_border",
                 "xme_mc_luma_batch", "xme_mc_chroma_batch", "xme_dct_batch", "xme_idct_batch", "xme_quant_batch",
                 "xme_dequant_batch", "xme_residual_batch", "xme_server_create", "xme_server_submit",
                 "xme_server_wait", "xme_me_search_sym_batch", "xme_me_search_bid_batch",
                 "xme_inter_recon_batch", "xme_inter_recon_bi_batch", "xme_get_mvp_batch"):
        assert must in names


@pytest.mark.skipif(not os.path.exists(LIB), reason="library not built")
def test_every_declared_symbol_is_exported():
    missing = [n for n in declared() if n not in exported()]
    assert not missing, f"declared in include/xavs2_amd.h but not exported: {missing}"


@pytest.mark.skipif(not os.path.exists(LIB), reason="library not built")
def test_library_loads_and_binds_without_a_gpu():
    L = C.CDLL(LIB)
    for n in declared():
        assert hasattr(L, n), n
    L.xme_version.restype = C.c_char_p
    assert b"gfx950" in L.xme_version()
    from xavs2_amd import lib as X
    X.lib()          # every argtypes assignment resolves a real symbol


@pytest.mark.skipif(not os.path.exists(LIB), reason="library not built")
def test_no_cpu_fallback_in_the_product():
    """The product library does not link the oracle."""
    out = subprocess.run(["nm", "-D", LIB], capture_output=True, text=True, check=True).stdout
    assert "xo_" not in out
    deps = subprocess.run(["ldd", LIB], capture_output=True, text=True).stdout
    assert "liboracle" not in deps


LAYOUTS = {
    "xme_pu": "PU_DTYPE", "xme_result": "RESULT_DTYPE", "xme_umh_ctx": "UMH_CTX_DTYPE",
    "xme_sym_job": "SYM_JOB_DTYPE", "xme_sym_result": "SYM_RESULT_DTYPE", "xme_bid_job": "BID_JOB_DTYPE",
    "xme_mvp_job": "MVP_JOB_DTYPE", "xme_mvp_result": "MVP_RESULT_DTYPE",
    "xme_umh_job": "UMH_JOB_DTYPE", "xme_umh_state": "UMH_STATE_DTYPE",
    "xme_bi_job": "BI_JOB_DTYPE", "xme_bi_result": "BI_RESULT_DTYPE",
    "xme_dual_job": "DUAL_JOB_DTYPE", "xme_dual_result": "DUAL_RESULT_DTYPE",
}


def test_struct_layouts_match_the_header(tmp_path):
    """sizeof / offsetof of the header's records, compiled with gcc, against the
    numpy dtypes and ctypes structures that mirror them."""
    from xavs2_amd import lib as X
    src = ['#include <stdio.h>', '#include <stddef.h>', '#include "xavs2_amd.h"', 'int main(void) {']
    for cname, dname in LAYOUTS.items():
        dt = getattr(X, dname)
        src.append(f'printf("{cname} %zu\\n", sizeof({cname}));')
        # numpy field names mirror the header's, except the padding bytes
        for f in dt.names:
            if f in ("r0", "r1"):
                continue
            src.append(f'printf("{cname}.{f} %zu\\n", offsetof({cname}, {f}));')
    for cname, sname in (("xme_me_params", "MeParams"), ("xme_server_config", "ServerConfig"),
                         ("xme_ref_desc", "RefDesc")):
        src.append(f'printf("{cname} %zu\\n", sizeof({cname}));')
    src.append("return 0; }")
    c = tmp_path / "layout.c"
    c.write_text("\n".join(src))
    exe = tmp_path / "layout"
    inc = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "include")
    subprocess.run(["gcc", "-std=c11", "-I", inc, str(c), "-o", str(exe)], check=True)
    got = dict(line.rsplit(" ", 1) for line in subprocess.run([str(exe)], capture_output=True, text=True,
                                                               check=True).stdout.splitlines())
    for cname, dname in LAYOUTS.items():
        dt = getattr(X, dname)
        assert int(got[cname]) == dt.itemsize, cname
        for f in dt.names:
            if f not in ("r0", "r1"):
                assert int(got[f"{cname}.{f}"]) == dt.fields[f][1], (cname, f)
    import ctypes
    for cname, sname in (("xme_me_params", "MeParams"), ("xme_server_config", "ServerConfig"),
                         ("xme_ref_desc", "RefDesc")):
        assert int(got[cname]) == ctypes.sizeof(getattr(X, sname)), cname


@pytest.mark.skipif(not os.path.exists(LIB), reason="library not built")
def test_chroma_qp_is_host_only_and_matches_the_oracle():
    """xme_chroma_qp is a table lookup on the host: it runs without a GPU."""
    from oracle import oracle as O
    L = C.CDLL(LIB)
    for q in range(-10, 81):
        assert L.xme_chroma_qp(q) == O.chroma_qp(q), q


@pytest.mark.skipif(not os.path.exists(LIB), reason="library not built")
def test_reconstruction_and_mc_arguments_are_checked_on_the_host():
    """The checks of the prediction-picture, chroma and bi-predicted MC entry points refuse
    bad sizes, QPs and missing buffers, and accept empty batches, before any device call."""
    from xavs2_amd import lib as X
    L = X.lib()
    p = C.c_void_p(4096)          # never dereferenced: every call below fails or returns early
    for n, qp in ((12, 30), (8, -1), (8, 80), (16 | X.XME_TU_DCT_HALF, 30), (128, 30)):
        assert L.xme_recon_pred_batch(n, qp, p, 64, p, 64, 64, 64, p, 1, p, p, 64, p, p, None) == X.XME_ERR_ARG
        assert L.xme_recon_chroma_batch(n, qp, p, p, 64, p, p, 64, 64, 64, p, 1, p, p, p, 64, p, p,
                                        None) == X.XME_ERR_ARG
    assert L.xme_recon_pred_batch(64 | X.XME_TU_DCT_HALF, 30, None, 64, None, 64, 64, 64, None, 0, None, None, 64,
                                  None, None, None) == X.XME_OK
    assert L.xme_recon_pred_batch(8, 30, None, 64, p, 64, 64, 64, p, 1, p, p, 64, p, p, None) == X.XME_ERR_ARG
    assert L.xme_recon_chroma_batch(8, 30, p, None, 64, p, p, 64, 64, 64, p, 1, p, p, p, 64, p, p,
                                    None) == X.XME_ERR_ARG
    # misaligned pictures, strides or levels
    odd = C.c_void_p(4097)
    for args in ((odd, 64, p, 64, p), (p, 66, p, 64, p), (p, 64, odd, 64, p), (p, 64, p, 65, p),
                 (p, 64, p, 64, C.c_void_p(4104))):
        cur, ic, pred, ip, lev = args
        assert L.xme_recon_pred_batch(8, 30, cur, ic, pred, ip, 64, 64, p, 1, lev, p, 64, p, p,
                                      None) == X.XME_ERR_ARG
    assert L.xme_mc_chroma_bi_batch(p, p, p, p, 64, None, 1, 64, 64, p, p, 64, p, None) == X.XME_ERR_ARG
    assert L.xme_mc_chroma_bi_batch(p, p, p, p, 64, p, -1, 64, 64, p, p, 64, p, None) == X.XME_ERR_ARG
    assert L.xme_mc_chroma_bi_batch(None, None, None, None, 64, None, 0, 64, 64, None, None, 64, None,
                                    None) == X.XME_OK
    r1, r2 = X.RefDesc(), X.RefDesc()
    r1.planes[0], r2.planes[0], r1.stride, r2.stride = 4096, 4096, 256, 320
    assert L.xme_mc_luma_bi_batch(C.byref(r1), C.byref(r2), p, 1, 64, 64, p, 64, p, None) == X.XME_ERR_ARG
    r2.stride = 256
    assert L.xme_mc_luma_bi_batch(C.byref(r1), C.byref(r2), None, 0, 64, 64, None, 64, None, None) == X.XME_OK
