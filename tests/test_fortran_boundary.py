"""The Fortran side of the boundary (include/scone_gpu_mod.f90, include/scone_gpu_adapter.f90),
compiled with the image's Fortran compiler (amdflang):
  - every bind(C) type has the size and every component the offset of its C struct in
    include/scone_gpu.h (gcc probe against an amdflang probe of the same list);
  - every entry point the header declares has an interface in the module;
  - the adapter converts a particleState's components to the C record and back (logical isMG
    as int32), and maps the transportOperatorGPU tracking names;
  - the module links against libscone_amd.so and its calls return the library's error codes."""
import os
import re
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
INC = os.path.join(ROOT, "include")

# struct -> components (C names; the Fortran components carry the same names)
LAYOUT = {
    "scn_particle_state": ["wgt", "r", "dir", "E", "G", "isMG", "type", "pad0", "time", "lambda", "matIdx", "cellIdx",
                           "uniqueID", "collisionN", "broodID", "pad1"],
    "scn_box": ["n_axes", "axes", "bc", "origin", "halfwidth"],
    "scn_surface": ["type", "axis", "tol", "c"],
    "scn_universe": ["type", "n_local", "origin", "border", "size", "out_local", "offset_mode", "offset_map", "pitch",
                     "corner", "a_bar", "outline_hw", "n_annuli", "annuli", "n_cells", "cells", "border_surf",
                     "rotated", "global", "check_overlap", "rot"],
    "scn_mg_data": ["n_mat", "n_groups", "xs", "P0", "prod", "chi", "fissile", "majorant", "collision_xs", "P1"],
    "scn_map1d": ["type", "axis", "nbins", "grid_type", "data"],
    "scn_clerk": ["gtop", "glow", "n_maps", "handle_virtual", "maps", "multi", "width", "resp", "kind", "mem_addr",
                  "size"],
    "scn_problem": ["transport", "st_cache", "n_universes", "root_uni", "universes", "n_boxes", "border_box", "boxes",
                    "graph_size", "graph_fill", "graph_id", "n_real", "real_data", "n_int", "int_data", "mg",
                    "n_clerks", "clerks", "n_bins", "norm_bin", "norm_value", "ht_cutoff", "n_surfaces", "surfaces",
                    "collision", "max_split", "ww_n", "fixed_point_tally", "ww_map", "ww_lower", "ww_upper", "ww_survival"],
    "scn_eigen_settings": ["pop", "n_inactive", "n_active", "inactive_tally", "seed", "keff0", "events_per_launch",
                           "block_size", "cycle_graph", "refill", "select_window"],
    "scn_eigen_result": ["k_analog", "k_implicit", "imp", "k_history", "cycle_time", "bins_active", "bins_inactive",
                         "transport_time", "n_collisions", "n_events", "n_histories"],
    "scn_source": ["type", "isotropic", "G", "mat", "n_probG", "pad", "r", "dir", "bottom", "top", "t_low", "t_high",
                   "probG"],
    "scn_fixed_request": ["problem", "source", "pop", "n_batches", "seed", "stop_bin", "stop_rel", "record", "buffer",
                          "window", "adapt"],
    "scn_fixed_result": ["state", "n_batches", "n_bins", "n_histories", "n_collisions", "n_segments"],
    "scn_batch_settings": ["transport", "max_slots", "hbm_budget", "tile", "k0", "n_launch", "use_graph", "refill",
                           "pad0"],
    "scn_stats": ["transport_kernel_ms", "transport_launches", "n_histories", "n_segments", "n_collisions", "n_sites",
                  "n_scores", "n_cycles", "cycle_ms"],
}


def _fc():
    fc = shutil.which("amdflang") or "/opt/rocm/bin/amdflang"
    if not os.path.exists(fc):
        pytest.skip("no Fortran compiler in this image")
    return fc


def _c_layout(tmp):
    lines = ["#include <stdio.h>", "#include <stddef.h>", '#include "scone_gpu.h"', "int main(void) {"]
    for st, fields in LAYOUT.items():
        lines.append(f'  printf("%zu", sizeof({st}));')
        for f in fields:
            lines.append(f'  printf(" %zu", offsetof({st}, {f}));')
        lines.append('  printf("\\n");')
    lines.append("  return 0; }")
    src = os.path.join(tmp, "probe.c")
    with open(src, "w") as fh:
        fh.write("\n".join(lines))
    exe = os.path.join(tmp, "probe_c")
    subprocess.run(["gcc", "-I", INC, src, "-o", exe], check=True)
    out = subprocess.run([exe], capture_output=True, text=True, check=True).stdout.strip().split("\n")
    return [[int(x) for x in l.split()] for l in out]


def _f_layout(tmp, fc):
    body = ["program probe", "  use iso_c_binding", "  use scone_gpu_mod", "  implicit none"]
    for i, st in enumerate(LAYOUT):
        body.append(f"  type({st}), target :: v{i}")
    body.append("  integer(c_intptr_t) :: base")
    body.append("  character(len=4096) :: line")
    for i, (st, fields) in enumerate(LAYOUT.items()):
        body.append(f"  base = transfer(c_loc(v{i}), base)")
        body.append(f"  write(line, '(I0)') c_sizeof(v{i})")
        for f in fields:
            body.append(f"  write(line, '(A,1X,I0)') trim(line), transfer(c_loc(v{i}%{f}), base) - base")
        body.append("  print '(A)', trim(line)")
    body.append("end program probe")
    src = os.path.join(tmp, "probe.f90")
    with open(src, "w") as fh:
        fh.write("\n".join(body))
    exe = os.path.join(tmp, "probe_f")
    subprocess.run([fc, "-c", os.path.join(INC, "scone_gpu_mod.f90"), "-J", tmp, "-o", os.path.join(tmp, "mod.o")],
                   check=True, cwd=tmp)
    subprocess.run([fc, "-I", tmp, src, os.path.join(tmp, "mod.o"), "-o", exe], check=True, cwd=tmp)
    out = subprocess.run([exe], capture_output=True, text=True, check=True).stdout.strip().split("\n")
    return [[int(x) for x in l.split()] for l in out]


def test_every_bind_c_type_matches_its_c_struct(tmp_path):
    fc = _fc()
    c = _c_layout(str(tmp_path))
    f = _f_layout(str(tmp_path), fc)
    for (st, fields), cl, fl in zip(LAYOUT.items(), c, f):
        assert cl == fl, f"{st}: C {cl} vs Fortran {fl} (size, then {fields})"


def test_every_declared_entry_point_has_a_fortran_interface():
    header = open(os.path.join(INC, "scone_gpu.h")).read()
    declared = set(re.findall(r"^\w[\w\s\*]*?\b(scn_\w+)\(", header, flags=re.M))
    mod = open(os.path.join(INC, "scone_gpu_mod.f90")).read()
    bound = set(re.findall(r'bind\(C, name="(scn_\w+)"\)', mod))
    assert declared == bound, (declared - bound, bound - declared)
    c_structs = set(re.findall(r"^typedef struct (scn_\w+)", header, flags=re.M))
    f_types = set(re.findall(r"type, public, bind\(C\) :: (scn_\w+)", mod))
    assert c_structs == f_types == set(LAYOUT), (c_structs ^ f_types)


ADAPTER_DRIVER = """
program drive
  use iso_c_binding
  use scone_gpu_mod
  use scone_gpu_adapter
  implicit none
  type(scn_particle_state) :: c
  real(c_double) :: wgt, r(3), dir(3), E, time, lambda
  integer(c_int32_t) :: G, typ, matIdx, cellIdx, uniqueID, collisionN, broodID, n_cu
  logical :: isMG
  type(c_ptr) :: eng
  integer(c_int) :: rc
  character(kind=c_char) :: name(64)
  call state_toC(0.5_c_double, [1.0_c_double, -2.0_c_double, 3.5_c_double], [0.0_c_double, 0.6_c_double, &
                 0.8_c_double], 0.0_c_double, 4_c_int32_t, .true., 1_c_int32_t, 2.25_c_double, 7.0_c_double, &
                 3_c_int32_t, 0_c_int32_t, 17_c_int32_t, 5_c_int32_t, 42_c_int32_t, c)
  print '(I0,1X,I0,1X,I0,1X,I0)', c % isMG, c % G, c % broodID, c % uniqueID
  call state_fromC(c, wgt, r, dir, E, G, isMG, typ, time, lambda, matIdx, cellIdx, uniqueID, collisionN, broodID)
  print '(L1,1X,F5.2,1X,F5.2,1X,F5.2,1X,I0,1X,I0)', isMG, wgt, r(3), time, collisionN, matIdx
  c % isMG = 0
  call state_fromC(c, wgt, r, dir, E, G, isMG, typ, time, lambda, matIdx, cellIdx, uniqueID, collisionN, broodID)
  print '(L1)', isMG
  print '(I0,1X,I0,1X,I0,1X,I0,1X,I0)', gpu_tracking('DT'), gpu_tracking('transportOperatorST'), &
        gpu_tracking(' HT'), gpu_tracking('MC'), gpu_tracking('transportOperatorDT')
  rc = scn_create(c_null_ptr, 10_c_int32_t, eng)
  print '(I0,1X,I0)', rc, len(gpu_message(scn_last_error(c_null_ptr)))
  print '(A)', gpu_message(scn_last_error(c_null_ptr))
  print '(I0)', len(gpu_message(c_null_ptr))
  rc = scn_device_info(name, 64_c_int32_t, n_cu)
  print '(I0)', rc
end program drive
"""


def test_adapter_converts_particle_states_and_names(tmp_path):
    fc = _fc()
    work = str(tmp_path)
    with open(os.path.join(work, "drive.f90"), "w") as fh:
        fh.write(ADAPTER_DRIVER)
    lib_dir = os.path.join(ROOT, "scone_amd")
    if not os.path.exists(os.path.join(lib_dir, "libscone_amd.so")):
        pytest.skip("libscone_amd.so is not built")
    for name in ("scone_gpu_mod", "scone_gpu_adapter"):
        subprocess.run([fc, "-c", os.path.join(INC, name + ".f90"), "-J", work, "-o", os.path.join(work, name + ".o")],
                       check=True, cwd=work)
    subprocess.run([fc, "-I", work, os.path.join(work, "drive.f90"), os.path.join(work, "scone_gpu_mod.o"),
                    os.path.join(work, "scone_gpu_adapter.o"), "-L", lib_dir, "-lscone_amd", f"-Wl,-rpath,{lib_dir}",
                    "-o", os.path.join(work, "drive")], check=True, cwd=work)
    out = subprocess.run([os.path.join(work, "drive")], capture_output=True, text=True, check=True).stdout
    lines = out.strip().split("\n")
    assert lines[0].split() == ["1", "4", "42", "17"]
    assert lines[1].split() == ["T", "0.50", "3.50", "2.25", "5", "3"]
    assert lines[2].strip() == "F"
    assert lines[3].split() == ["2", "1", "3", "0", "2"]
    rc, n = (int(x) for x in lines[4].split())
    assert rc != 0 and n > 0                      # no GPU here, or no problem tables
    assert len(lines[5].strip()) == n
    assert lines[6].strip() == "0"
