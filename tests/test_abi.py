"""The C-ABI library on the CPU: loads, exports every declared symbol, has the
reference's struct layouts, defaults, validation rules and messages
(src/common/mr_config.cpp:74-238 of the reference). No GPU work is issued."""

import ctypes
import re
import subprocess
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]
HEADER = ROOT / "include" / "mr_aha" / "mr_aha.h"


@pytest.fixture(scope="module")
def L():
    from mr_aha_amd import _lib
    return _lib


def declared_symbols():
    text = HEADER.read_text() + (HEADER.parent / "mr_aha_ext.h").read_text()
    return sorted(set(re.findall(r"MR_EXPORT\s+[\w\s\*]+?\b(mr_\w+)\s*\(", text)))


def test_exports_every_declared_symbol(L):
    syms = declared_symbols()
    # 25 reference entry points + mr_version + 14 extensions (profiling,
    # pipeline, fused-form and tiling queries, autotuner, kernel format, CG and
    # its allocation-free variant, the streaming copy of the bandwidth
    # calibration)
    assert len(syms) == 40
    for s in syms:
        assert hasattr(L.lib, s), s
    out = subprocess.run(["nm", "-D", "--defined-only", str(L.LIB_PATH)], capture_output=True,
                         text=True, check=True).stdout
    exported = {ln.split()[-1] for ln in out.splitlines() if " T " in ln}
    assert set(syms) <= exported


LAYOUT_C = r"""
#include <stdio.h>
#include <stddef.h>
#include "mr_aha/mr_aha.h"
#define F(T, f) printf(#T "." #f " %zu\n", offsetof(T, f));
int main(void) {
  printf("mr_config %zu\nmr_op_inputs %zu\nmr_op_options %zu\nmr_op_info_t %zu\nmr_complex %zu %zu\n",
         sizeof(mr_config), sizeof(mr_op_inputs), sizeof(mr_op_options), sizeof(mr_op_info_t),
         sizeof(mr_complex), _Alignof(mr_complex));
  F(mr_config, num_masked) F(mr_config, n_coils) F(mr_config, eps) F(mr_config, stream)
  F(mr_config, mem_fraction) F(mr_config, mask_radius)
  F(mr_op_inputs, Ntrajs) F(mr_op_inputs, d_phi2_toep)
  F(mr_op_options, devices) F(mr_op_options, nufft_upsampfac) F(mr_op_options, nufft_gpu_method)
  F(mr_op_info_t, n_split_coils) F(mr_op_info_t, peer_bandwidth_gbs) F(mr_op_info_t, create_ms)
  return 0;
}
"""


def test_struct_layouts(tmp_path, L):
    src = tmp_path / "layout.c"
    src.write_text(LAYOUT_C)
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-std=c11", f"-I{ROOT / 'include'}", str(src), "-o", str(exe)], check=True)
    vals = {}
    for ln in subprocess.run([str(exe)], capture_output=True, text=True, check=True).stdout.splitlines():
        k, *v = ln.split()
        vals[k] = [int(t) for t in v]
    # sizes of the reference's header on LP64
    assert vals["mr_config"] == [104] and vals["mr_op_inputs"] == [56]
    assert vals["mr_op_options"] == [48] and vals["mr_op_info_t"] == [320]
    assert vals["mr_complex"] == [8, 8]
    assert vals["mr_config.num_masked"] == [48] and vals["mr_config.eps"] == [64]
    assert vals["mr_config.stream"] == [80]
    assert ctypes.sizeof(L.MRConfig) == 104 and ctypes.sizeof(L.MROpInputs) == 56
    assert ctypes.sizeof(L.MROpOptions) == 48 and ctypes.sizeof(L.MROpInfo) == 320
    for key, v in vals.items():
        if "." in key:
            off = v[0]
            st, field = key.split(".")
            cls = {"mr_config": L.MRConfig, "mr_op_inputs": L.MROpInputs,
                   "mr_op_options": L.MROpOptions, "mr_op_info_t": L.MROpInfo}[st]
            assert getattr(cls, field).offset == off, key


def test_layout_equals_reference_header(tmp_path):
    """Every struct's size and alignment, every field's offset and size and
    every enumerator / MR_ constant of this header equal the reference
    header's (built from the reference checkout with gcc by
    oracle/abi_layout.py, kept in oracle/_ref/abi_ref.txt)."""
    from oracle import abi_layout
    ref = abi_layout.reference_layout()
    if ref is None:
        pytest.skip("neither the reference header nor oracle/_ref/abi_ref.txt is present")
    ours = abi_layout.run(ROOT / "include", tmp_path)
    assert len(ours.splitlines()) > 70
    assert ours.splitlines() == ref.splitlines()


def test_identity_and_status(L):
    assert L.backend() == "cuda"
    assert L.lib.mr_version() == 10000
    assert L.lib.mr_status_string(0) == b"success"
    assert L.lib.mr_status_string(-3) == b"GPU memory allocation failed"
    assert L.lib.mr_status_string(-7) == b"unsupported configuration"
    assert L.lib.mr_status_string(42) == b"unknown status"


def test_defaults_and_sizes(L):
    c = L.MRConfig()
    assert L.lib.mr_config_defaults(ctypes.byref(c), 2) == 0
    assert (c.traj_type, c.n_spatial, list(c.n_dim), list(c.phi2_dim), c.L) == (2, 3, [1, 1, 1], [1, 1, 1], 1)
    assert (c.triangulation, c.is_masked, c.num_masked, c.n_coils) == (0, 0, 0, 1)
    assert (c.eps, c.gpu_device_id, c.stream, c.mem_fraction, c.verbose, c.mask_radius) == \
        (1e-6, 0, None, 0.5, 0, 2)
    assert L.lib.mr_config_defaults(ctypes.byref(c), 0) == 0 and c.n_spatial == 2
    assert L.lib.mr_config_defaults(None, 0) == -1
    assert L.lib.mr_num_tri(16) == 136 and L.lib.mr_num_tri(0) == 0 and L.lib.mr_num_tri(-3) == 0
    from mr_aha_amd import make_config
    cfg = make_config("stack", (320, 320, 56), 16, n_coils=12)
    assert list(cfg.phi2_dim) == [640, 640, 56]
    assert L.lib.mr_phi2_toep_size(ctypes.byref(cfg)) == 3_119_513_600
    assert L.lib.mr_pad_size(ctypes.byref(cfg)) == 22_937_600
    assert L.lib.mr_image_size(ctypes.byref(cfg)) == 5_734_400
    assert L.lib.mr_x_size(ctypes.byref(cfg)) == 5_734_400 * 16
    assert L.lib.mr_se_size(ctypes.byref(cfg)) == 5_734_400 * 12
    assert L.lib.mr_pad_size(None) == 0


def _base(L, tt=0):
    c = L.MRConfig()
    L.lib.mr_config_defaults(ctypes.byref(c), tt)
    c.n_dim[0], c.n_dim[1] = 8, 6
    c.phi2_dim[0], c.phi2_dim[1] = 16, 12
    return c


@pytest.mark.parametrize("edit,status,msg", [
    (lambda c: setattr(c, "traj_type", 7), -2, "traj_type must be one of"),
    (lambda c: setattr(c, "n_spatial", 3), -2, "n_spatial is 3 but trajectory type requires 2"),
    (lambda c: c.n_dim.__setitem__(0, 0), -2, "n_dim[0] = 0 must be >= 1"),
    (lambda c: c.phi2_dim.__setitem__(1, 5), -2, "phi2_dim[1] = 5 must not be smaller than n_dim[1] = 6"),
    (lambda c: c.n_dim.__setitem__(2, 2), -2, "n_dim[2] and phi2_dim[2] must be 1 for a 2D trajectory"),
    (lambda c: setattr(c, "L", 0), -2, "L = 0 must be >= 1"),
    (lambda c: setattr(c, "n_coils", 0), -2, "n_coils = 0 must be >= 1"),
    (lambda c: setattr(c, "triangulation", 2), -2, "triangulation must be MR_TRI_LOWER or MR_TRI_UPPER"),
    (lambda c: setattr(c, "eps", 1.0), -2, "eps = 1 must lie in (0, 1)"),
    (lambda c: setattr(c, "mask_radius", -1), -2, "mask_radius = -1 must be >= 0"),
    (lambda c: (setattr(c, "is_masked", 1), setattr(c, "num_masked", 0)), -2,
     "num_masked = 0 must be > 0 when is_masked is set"),
    (lambda c: (setattr(c, "is_masked", 1), setattr(c, "num_masked", 10**6)), -2,
     "num_masked = 1000000 exceeds the upsampled grid size 192"),
])
def test_validation_rules_and_messages(L, edit, status, msg):
    c = _base(L)
    assert L.lib.mr_config_validate(ctypes.byref(c)) == 0
    assert L.lib.mr_last_error() == b""
    edit(c)
    assert L.lib.mr_config_validate(ctypes.byref(c)) == status
    assert msg in L.lib.mr_last_error().decode()
    # the next call clears the message
    assert L.lib.mr_config_defaults(ctypes.byref(c), 0) == 0
    assert L.lib.mr_last_error() == b""


def test_stack_mz_rule(L):
    c = L.MRConfig()
    L.lib.mr_config_defaults(ctypes.byref(c), 2)
    for a, (n, m) in enumerate([(8, 16), (6, 12), (4, 8)]):
        c.n_dim[a], c.phi2_dim[a] = n, m
    assert L.lib.mr_config_validate(ctypes.byref(c)) == -2
    assert b"for MR_TRAJ_STACK, phi2_dim[2] (Mz = 8) must equal n_dim[2] (Nz = 4)" in L.lib.mr_last_error()


def test_null_argument_errors(L):
    assert L.lib.mr_config_validate(None) == -1
    assert b"cfg is NULL" in L.lib.mr_last_error()
    assert L.lib.mr_aha_plan_create(None, None) == -1
    assert L.lib.mr_op_create(None, None, None, None) == -1
    assert b"op_out is NULL" in L.lib.mr_last_error()
    assert L.lib.mr_op_destroy(None) == 0 and L.lib.mr_aha_plan_destroy(None) == 0
    assert L.lib.mr_op_info(None, None) == -1
    c = _base(L)
    assert L.lib.mr_calc_toep(ctypes.byref(c), None, None, 1, 1, None, None) == -1
    assert b"mr_calc_toep: d_phi2 is NULL" in L.lib.mr_last_error()


def test_make_config_python_rules():
    from mr_aha_amd import MRError, make_config
    with pytest.raises(ValueError, match="n_dim must have 3"):
        make_config("stack", (4, 4), 2)
    with pytest.raises(ValueError, match="traj_type"):
        make_config("4d", (4, 4), 2)
    with pytest.raises(MRError, match="eps"):
        make_config("2D", (4, 4), 2, eps=2.0)
    cfg = make_config("3D", (5, 6, 7), 3, is_masked=True)  # num_masked filled in later
    assert cfg.is_masked == 1 and list(cfg.phi2_dim) == [10, 12, 14]
