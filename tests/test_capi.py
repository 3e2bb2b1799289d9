"""CPU-side checks of the drop-in boundary: the HIP library loads, exports every
symbol include/wtv_gpupr.h declares, and rejects bad arguments before touching a
device; the pybind module keeps the reference's Python signature."""
import ctypes as C
import inspect
import os
import re

import pytest

import wtv_capi as W

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    src = open(os.path.join(ROOT, "include", "wtv_gpupr.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(wtv_[a-z_]+)\s*\(", src)))


def test_header_symbols_exported():
    names = _declared()
    assert set(names) == set(W.EXPORTS)
    lib = W.lib()
    for nm in names:
        assert hasattr(lib, nm), nm


def test_params_defaults_match_reference():
    p = W.params()
    # wtivo_gpupr_bindings.cpp:587-590 defaults
    assert p.global_relabel_period == 1024 and p.local_steps == 8 and p.max_rounds == 2000000
    assert W.lib().wtv_version().startswith(b"wtv_gpupr")


def test_struct_layouts_match_c_header(tmp_path):
    """Every ctypes mirror has the size and field offsets that gcc gives the
    structs of include/wtv_gpupr.h (a field missing from either side fails to
    compile or to match)."""
    import subprocess
    pairs = [(W.WtvStats, "wtv_stats"), (W.WtvParams, "wtv_params"), (W.WtvCutResult, "wtv_cut_result"),
             (W.WtvServerConfig, "wtv_server_config"), (W.WtvRequest, "wtv_request"), (W.WtvResult, "wtv_result"),
             (W.WtvBuildInfo, "wtv_build_info")]
    lines, want = [], []
    for cls, cname in pairs:
        lines.append(f'printf("%zu\\n", sizeof({cname}));')
        want.append(C.sizeof(cls))
        for f in cls._fields_:
            lines.append(f'printf("%zu\\n", offsetof({cname}, {f[0]}));')
            want.append(getattr(cls, f[0]).offset)
    src = tmp_path / "layout.c"
    src.write_text("#include <stddef.h>\n#include <stdio.h>\n#include \"wtv_gpupr.h\"\nint main(void) {\n"
                   + "\n".join(lines) + "\nreturn 0;\n}\n")
    exe = tmp_path / "layout"
    inc = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "include")
    subprocess.run(["gcc", "-std=c99", "-I", inc, str(src), "-o", str(exe)], check=True)
    got = [int(x) for x in subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout.split()]
    assert got == want


def test_zero_nodes_and_null_pointers_without_device():
    st = W.pr_solve(0, 0, 0, 0, 0, 0)  # n == 0 returns zero stats (gpu_push_relabel_fast.cu:418)
    assert st.raw_flow == 0 and st.rounds == 0
    with pytest.raises(W.WtvError) as e:
        W.pr_solve(5, 0, 0, 0, 0, 0)
    assert e.value.code == W.WTV_EINVAL and "null graph pointer" in str(e.value)


def test_graph_cut_null_args_rejected():
    r = W.WtvCutResult()
    rc = W.lib().wtv_graph_cut_fast(None, 0, None, None, 10, None, 10.0, 1, 1024, 0, 8, None, C.byref(r), None)
    assert rc == W.WTV_EINVAL


def test_pybind_module_signature():
    import wtivo_gpupr as G
    for fn in (G.graph_cut_fast, G.graph_cut_gpupr_fast_v630):
        doc = fn.__doc__
        assert "vertices" in doc and "tets_cuda" in doc and "neighbors_cuda" in doc and "labels" in doc
        import re as _re
        got = dict(_re.findall(r"(\w+): [^=,]*= (\d+)", doc))
        assert got == {"threads": "16", "global_relabel_period": "1024", "max_rounds": "2000000",
                       "local_steps": "8"}
        order = _re.findall(r"(\w+): ", doc.split("(", 1)[1])
        assert order == ["vertices", "tets_cuda", "neighbors_cuda", "labels", "fill_T", "threads",
                         "global_relabel_period", "max_rounds", "local_steps"]


def test_pybind_rejects_non_cuda_like_reference():
    import numpy as np
    import torch
    import wtivo_gpupr as G
    V = np.zeros((4, 3))
    t = torch.zeros((1, 4), dtype=torch.int32)
    with pytest.raises(RuntimeError, match="tets/neighbors must be CUDA tensors"):
        G.graph_cut_fast(V, t, t, np.ones(1, np.uint8), 10.0)
    with pytest.raises(RuntimeError, match="vertices must be float64 Nx3"):
        G.graph_cut_fast(np.zeros((4, 2)), t, t, np.ones(1, np.uint8), 10.0)


def test_request_footprint_counts_1024_node_tiles():
    f = W.lib().wtv_request_footprint
    assert f(0) == 0 and f(1) == f(1024) > 0 and f(1025) == 2 * f(1024)
    assert f(3000) == 3 * f(1) and f(10 ** 6) == 977 * f(1)


def test_topology_entry_points_reject_bad_arguments_without_device():
    L = W.lib()
    assert L.wtv_tet_neighbors(None, 0, 1, None, None) == W.WTV_OK           # N == 0: nothing to do
    assert L.wtv_tet_neighbors(None, 10, 5, None, None) == W.WTV_EINVAL
    buf = (C.c_int32 * 16)()
    assert L.wtv_tet_neighbors(buf, 1 << 31, 5, buf, None) == W.WTV_EINVAL   # ids must fit int32
    assert L.wtv_tet_neighbors(buf, 4, 0, buf, None) == W.WTV_EINVAL         # no vertices
    assert L.wtv_topology_upload(None, 0, 4, 1, None, 0, None) == W.WTV_OK
    assert L.wtv_topology_upload(None, 5, 4, 1, None, 0, None) == W.WTV_EINVAL
    assert L.wtv_topology_upload(buf, 4, 0, 1, buf, 0, None) == W.WTV_EINVAL  # every row aliased
    assert "row stride" in W.lib().wtv_last_error().decode()


def test_topology_to_cuda_chunked_shape_check_like_reference():
    import numpy as np
    import wtv_topology as T
    with pytest.raises(RuntimeError, match=r"tets: expected Nx4 topology array, got \(3, 3\)"):
        T.topology_to_cuda_chunked(np.zeros((3, 3), np.int32), "tets")


def test_udf_entry_points_reject_bad_arguments_without_device():
    L = W.lib()
    h = C.c_void_p(0)
    assert L.wtv_udf_create(None, None, 0, 1.0, 0.1, C.byref(h), None) == W.WTV_EINVAL
    assert "grid is empty" in L.wtv_last_error().decode()
    o = (C.c_int32 * 3)(0, 0, 4)
    v = (C.c_float * 512)()
    assert L.wtv_udf_create(o, v, 1, 1.0, 0.1, C.byref(h), None) == W.WTV_EINVAL    # origin not a multiple of 8
    assert L.wtv_udf_create(o, v, 1, 1.0, 0.1, None, None) == W.WTV_EINVAL          # no handle pointer
    b0 = (C.c_double * 3)(0, 0, 0)
    b1 = (C.c_double * 3)(1, 1, 1)
    assert L.wtv_udf_sample_tet_labels(None, None, 0, 0, None, 0, b0, b1, 0.0, 0.5, None, 0, None, None) == W.WTV_EINVAL
    assert "grid is empty" in L.wtv_last_error().decode()


def test_resident_cut_rejects_null_arguments_without_device():
    L = W.lib()
    r = W.WtvCutResult()
    assert L.wtv_graph_cut_fast_resident(None, 0, None, None, 10, None, 10.0, 1024, 0, 8, None, C.byref(r),
                                         None) == W.WTV_EINVAL
    assert L.wtv_graph_cut_fast_resident(None, 0, None, None, 0, None, 10.0, 1024, 0, 8, None, C.byref(r),
                                         None) == W.WTV_OK          # N == 0: nothing to do


def test_flow_mode_zero_nodes_and_null_pointers_without_device():
    st, ret = W.pr_solve_flow(0, 0, 0, 0, 0, 0)
    assert st.raw_flow == 0 and ret == 0
    with pytest.raises(W.WtvError) as e:
        W.pr_solve_flow(5, 0, 0, 0, 0, 0)
    assert e.value.code == W.WTV_EINVAL
