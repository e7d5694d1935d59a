"""C-ABI boundary checks that run without a GPU."""
import ctypes as C
import os
import re

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "gvdbx", "gvdbx_c.h")


def _declared():
    txt = open(HEADER).read()
    return sorted(set(re.findall(r"^\s*(?:const\s+)?\w+\*?\s+\*?(gvdbx_\w+)\s*\(", txt, re.M)))


def test_library_exports_every_declared_symbol():
    import gvdbx
    L = gvdbx.lib()
    names = _declared()
    assert len(names) >= 30
    for n in names:
        assert hasattr(L, n), f"{n} declared in gvdbx_c.h but not exported"
    # the ctypes signature table covers the whole header
    assert set(names) == set(gvdbx.exported_names())


def test_no_cpu_fallback_symbols():
    import gvdbx
    out = os.popen(f"nm -D {gvdbx._lib.LIB_PATH}").read()
    assert "or_rebuild_topology" not in out      # the oracle is never linked into the product


def test_create_without_gpu_fails_cleanly():
    torch = pytest.importorskip("torch")
    if torch.cuda.device_count() > 0:
        pytest.skip("a GPU is present")
    import gvdbx
    h = C.c_void_p()
    code = gvdbx.lib().gvdbx_create(C.byref(h), 0)
    assert code == 6 and not h.value
    assert "device" in gvdbx.last_error()


def test_null_handle_errors():
    import gvdbx
    L = gvdbx.lib()
    assert L.gvdbx_rebuild_topology(None, 10, 1.0, 0.0, 0.0, 0.0) == 1
    assert "null" in gvdbx.last_error()
    assert L.gvdbx_destroy(None) == 1


def test_generators_deterministic_and_in_range():
    import gvdbx
    a = gvdbx.gen_box(7, 10000, (64, 64, 64), (576, 320, 576), 0.5)
    b = gvdbx.gen_box(7, 10000, (64, 64, 64), (576, 320, 576), 0.5)
    np.testing.assert_array_equal(a, b)
    assert a.min(0)[0] >= 63.5 and a.max(0)[1] <= 320.5
    s = gvdbx.gen_shell(8, 20000, (256, 256, 256), 180, 200)
    r = np.linalg.norm(s.astype(np.float64) - 256, axis=1)
    assert r.min() >= 180 - 1e-3 and r.max() <= 200 + 1e-3
    # uniform by volume: median radius = cbrt((180^3+200^3)/2)
    assert abs(np.median(r) - ((180 ** 3 + 200 ** 3) / 2) ** (1 / 3)) < 0.5


def test_cpp_facade_compiles_and_links(tmp_path):
    """The header-only VolumeGVDB facade compiles against the C-ABI and links libgvdbx.so."""
    import shutil
    import subprocess
    if shutil.which("g++") is None:
        pytest.skip("no g++")
    src = tmp_path / "facade.cpp"
    src.write_text(r'''
#include "gvdbx/VolumeGVDB.h"
#include <cstdio>
int main(int argc, char**) {
    if (argc > 5) {   // never executed here: only the call sequence is compiled and linked
        gvdbx::VolumeGVDB gvdb;
        gvdb.Configure(3, 3, 3, 3, 5);
        gvdb.SetChannelDefault(16, 16, 1);
        gvdb.AddChannel(0, gvdbx::T_FLOAT, 1);
        gvdbx::DataPtr pos, vel, clr;
        gvdb.SetDataGPU(pos, 10, nullptr, 0, 12);
        gvdb.SetPoints(pos, vel, clr);
        gvdb.RebuildTopology(10, 3.0f, gvdbx::Vector3DF(0, 0, 0));
        gvdb.FinishTopology(false, true);
        gvdb.UpdateAtlas();
        int len = 0;
        gvdb.InsertPointsSubcell(4, 10, 3.0f, gvdbx::Vector3DF(0, 0, 0), len);
        gvdb.GatherLevelSet(4, 10, 1.5f, gvdbx::Vector3DF(0, 0, 0), len, 0, 1);
        gvdb.UpdateApron(0, 3.0f);
        gvdbx_camera cam{};
        gvdbx_render_params rp{};
        gvdb.Render(GVDBX_SHADE_VOLUME, 0, cam, rp, nullptr, nullptr);
        gvdb.SetTuning("splat_chunk", 1024);
        gvdb.SplatCubic(10, 2.0f, gvdbx::Vector3DF(0, 0, 0));
        gvdb.RaycastEA(0, cam, rp, nullptr, nullptr, true);
        gvdb.RaycastEABackward(0, cam, rp, nullptr, nullptr, nullptr, nullptr, nullptr);
        gvdb.ApronAdjoint(nullptr);
        gvdb.SplatCubicBackward(10, 2.0f, gvdbx::Vector3DF(0, 0, 0), nullptr, nullptr);
        gvdb.SetPointsBatched(nullptr, 12, std::vector<int64_t>{0, 5, 10});
        std::vector<gvdbx_camera> cams(2);
        gvdb.RaycastEABatched(0, cams, rp, nullptr, nullptr, true);
        gvdb.RaycastEABackwardBatched(0, cams, rp, nullptr, nullptr, nullptr, nullptr, nullptr);
    }
    std::printf("%s\n", gvdbx_version());
    return 0;
}
''')
    import gvdbx
    libdir = os.path.dirname(gvdbx._lib.LIB_PATH)
    exe = tmp_path / "facade"
    cmd = ["g++", "-std=c++17", "-I", os.path.join(ROOT, "include"), str(src), "-L", libdir, "-lgvdbx",
           f"-Wl,-rpath,{libdir}", "-o", str(exe)]
    subprocess.run(cmd, check=True)
    out = subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout
    assert "gvdbx" in out
