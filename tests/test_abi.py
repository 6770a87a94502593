"""CPU checks of the C-ABI boundary: the HIP library loads, exports every
entry point include/lsdg.h declares, and the ctypes mirror has the C layout."""
import ctypes as C
import glob
import os
import re
import subprocess

import pytest

import lsdg

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HDRS = sorted(glob.glob(os.path.join(ROOT, "include", "*.h")))


def declared_functions():
    txt = "".join(open(h).read() for h in HDRS)
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(lsdg_[a-z0-9_]+)\s*\(", txt)))


def test_header_declares_entry_points():
    fns = declared_functions()
    for must in ("lsdg_se3_track", "lsdg_sim3_track", "lsdg_depth_update", "lsdg_depth_create_keyframe",
                 "lsdg_energy_grad", "lsdg_energy_grad_sim3", "lsdg_batch_track", "lsdg_batch_tune"):
        assert must in fns


def test_library_exports_every_declared_symbol():
    L = C.CDLL(lsdg.lib_path())
    missing = [f for f in declared_functions() if not hasattr(L, f)]
    assert not missing, missing
    lsdg.load()  # argtypes for every wrapped symbol resolve


def test_struct_layouts_match_c(tmp_path):
    src = tmp_path / "sizes.c"
    src.write_text('#include <stdio.h>\n#include <stddef.h>\n#include "lsdg_tune.h"\n'
                   'int main(void){printf("%zu %zu %zu %zu %zu %zu %zu %zu %zu %zu %zu %zu %zu %zu %zu %zu %zu %zu %zu\\n", '
                   'sizeof(lsdg_params), '
                   'sizeof(lsdg_track_stats), sizeof(lsdg_eval), sizeof(lsdg_hypotheses), sizeof(lsdg_camera), '
                   'sizeof(lsdg_job), sizeof(lsdg_job_result), offsetof(lsdg_params, stereo_max_steps), '
                   'sizeof(lsdg_point_dense), sizeof(lsdg_vertex), sizeof(lsdg_vbo_params), '
                   'sizeof(lsdg_constraint_candidate), sizeof(lsdg_constraint_edge), sizeof(lsdg_constraint_result), '
                   'offsetof(lsdg_constraint_result, frame_on_candidate), sizeof(lsdg_system_params), '
                   'sizeof(lsdg_system_frame), offsetof(lsdg_system_frame, stats), sizeof(lsdg_batch_config));'
                   'return 0;}\n')
    exe = tmp_path / "sizes"
    subprocess.check_call(["gcc", "-std=c99", "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe)])
    got = [int(v) for v in subprocess.check_output([str(exe)]).split()]
    want = [C.sizeof(lsdg.Params), C.sizeof(lsdg.TrackStats), C.sizeof(lsdg.Eval), C.sizeof(lsdg.Hypotheses),
            C.sizeof(lsdg.Camera), C.sizeof(lsdg.Job), C.sizeof(lsdg.JobResult), lsdg.Params.stereo_max_steps.offset,
            C.sizeof(lsdg.PointDense), C.sizeof(lsdg.Vertex), C.sizeof(lsdg.VboParams),
            C.sizeof(lsdg.ConstraintCandidate), C.sizeof(lsdg.ConstraintEdge), C.sizeof(lsdg.ConstraintResult),
            lsdg.ConstraintResult.frame_on_candidate.offset, C.sizeof(lsdg.SystemParams), C.sizeof(lsdg.SystemFrame),
            lsdg.SystemFrame.stats.offset, C.sizeof(lsdg.BatchConfig)]
    assert got == want


def test_default_params_match_oracle():
    import pyoracle
    a = lsdg.default_params().as_dict()
    b = pyoracle.default_params().as_dict()
    assert a == b


def test_product_does_not_import_oracle():
    pkg = os.path.join(ROOT, "lsd-slam-pangolin-gui_amd")
    for dirpath, _, files in os.walk(pkg):
        for f in files:
            if f.endswith((".py", ".cpp", ".hip", ".h", ".hpp")) and "_build" not in dirpath:
                txt = open(os.path.join(dirpath, f)).read()
                assert "pyoracle" not in txt and "liboracle" not in txt and "lsd_oracle" not in txt, f


def test_shim_classes_compile_and_link(tmp_path):
    """The lsd_slam shim (upstream class names and signatures over the C ABI)
    compiles as host C++ and links against liblsdg.so; no device call is
    made here."""
    src = tmp_path / "use_shim.cpp"
    src.write_text('#include "lsd_slam_shim.h"\n'
                   'int main() {\n'
                   '  using namespace lsd_slam;\n'
                   '  SE3Tracker* (*mk)(Context&) = [](Context& c) { return new SE3Tracker(c); };\n'
                   '  SE3 (SE3Tracker::*t1)(TrackingReference*, Frame*, const SE3&) = &SE3Tracker::trackFrame;\n'
                   '  SE3 (SE3Tracker::*t2)(TrackingReference*, const Frame::SharedPtr&, const SE3&) =\n'
                   '      &SE3Tracker::trackFrame;\n'
                   '  Sim3 (Sim3Tracker::*t3)(TrackingReference*, Frame*, const Sim3&, int, int) =\n'
                   '      &Sim3Tracker::trackFrameSim3;\n'
                   '  void (DepthMap::*u)(const std::deque<std::shared_ptr<Frame>>&) = &DepthMap::updateKeyframe;\n'
                   '  void (DepthMap::*k)(Frame*) = &DepthMap::createKeyFrame;\n'
                   '  lsdg_params p; lsdg_params_default(&p);\n'
                   '  return (mk && t1 && t2 && t3 && u && k && p.huber_d == 3.0f) ? 0 : 1;\n'
                   '}\n')
    exe = tmp_path / "use_shim"
    pkg = os.path.join(ROOT, "lsd-slam-pangolin-gui_amd")
    subprocess.check_call(["g++", "-std=c++17", "-Wall", "-Werror", "-I", os.path.join(pkg, "shim"), str(src),
                           "-o", str(exe), "-L", pkg, "-llsdg", "-Wl,-rpath," + pkg])
    assert subprocess.call([str(exe)]) == 0
