"""CPU-side checks of the C-ABI library: it loads, exports every function include/frehg_gpu.h
declares, reports its ABI version, and refuses to create a context without a GPU."""
import ctypes as C
import os
import subprocess

import pytest

from frehg_amd import abi, gpu


def test_library_exports_every_declared_symbol():
    lib = gpu.load()
    names = gpu.declared_symbols()
    assert len(names) >= 30
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def test_abi_version():
    assert gpu.load().frehg_abi_version() == abi.ABI_VERSION


def test_struct_layouts_match_the_c_header(tmp_path):
    """The ctypes mirror and the header agree on size and field offsets (checked by gcc)."""
    structs = {"frehg_soil_params": abi.SoilParams, "frehg_solver_config": abi.SolverConfig,
               "frehg_re_params": abi.ReParams, "frehg_swe_params": abi.SweParams,
               "frehg_coupling_params": abi.CouplingParams, "frehg_solve_stats": abi.SolveStats,
               "frehg_re_stats": abi.ReStats, "frehg_coupled_stats": abi.CoupledStats,
               "frehg_ens_stats": abi.EnsStats, "frehg_run_setup": abi.RunSetup, "frehg_run_stats": abi.RunStats,
               "frehg_checkpoint_info": abi.CheckpointInfo}
    lines = ["#include <stdio.h>", "#include <stddef.h>", '#include "frehg_gpu.h"', "int main(void){"]
    for cname, py in structs.items():
        lines.append(f'printf("{cname} %zu\\n", sizeof({cname}));')
        for fname, _ in py._fields_:
            lines.append(f'printf("{cname}.{fname} %zu\\n", offsetof({cname}, {fname}));')
    lines.append("return 0;}")
    src = tmp_path / "layout.c"
    src.write_text("\n".join(lines))
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-I", os.path.dirname(gpu.HEADER), str(src), "-o", str(exe)], check=True)
    out = dict(l.split() for l in subprocess.run([str(exe)], capture_output=True, text=True, check=True).stdout.splitlines())
    for cname, py in structs.items():
        assert int(out[cname]) == C.sizeof(py), cname
        for fname, _ in py._fields_:
            assert int(out[f"{cname}.{fname}"]) == getattr(py, fname).offset, f"{cname}.{fname}"


@pytest.mark.skipif(os.path.exists("/dev/kfd") and os.access("/dev/kfd", os.R_OK), reason="GPU present")
def test_create_without_gpu_fails_cleanly():
    h = C.c_void_p()
    rc = gpu.load().frehg_create(C.byref(h), 0)
    assert rc == abi.ERR_NO_DEVICE and not h.value
