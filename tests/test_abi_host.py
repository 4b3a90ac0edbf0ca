"""Host-side checks without a GPU: the C ABI library loads and exports every entry point the
header declares; the ctypes mirrors match the header's struct layouts; the dictionary reader and
the table compiler accept SCONE's input format."""
import ctypes as C
import os
import re
import subprocess

import numpy as np
import pytest

from scone_amd import configs, dictio, engine, tables as T

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "scone_gpu.h")


def declared_functions():
    text = open(HEADER).read()
    return sorted(set(re.findall(r"^\w[\w\s\*]*?\b(scn_\w+)\(", text, flags=re.M)))


def test_library_exports_every_declared_entry_point():
    L = engine.load()
    names = declared_functions()
    assert len(names) >= 12
    for n in names:
        assert hasattr(L, n), n
    assert sorted(engine.EXPORTS) == names


def test_struct_layouts_match_header(tmp_path):
    src = r'''
#include <stdio.h>
#include <stddef.h>
#include "scone_gpu.h"
int main(void) {
  printf("%zu %zu %zu %zu %zu %zu %zu %zu %zu %zu\n", sizeof(scn_particle_state), sizeof(scn_box),
         sizeof(scn_universe), sizeof(scn_mg_data), sizeof(scn_map1d), sizeof(scn_clerk), sizeof(scn_problem),
         sizeof(scn_eigen_settings), sizeof(scn_eigen_result), sizeof(scn_stats));
  printf("%zu %zu %zu\n", offsetof(scn_particle_state, time), offsetof(scn_particle_state, broodID),
         offsetof(scn_problem, mg));
  printf("%zu %zu %zu %zu %zu %zu\n", sizeof(scn_source), sizeof(scn_fixed_request), sizeof(scn_fixed_result),
         sizeof(scn_batch_settings), offsetof(scn_source, probG), offsetof(scn_fixed_request, stop_rel));
  printf("%zu %zu %zu %zu %zu %zu %zu %zu\n", sizeof(scn_surface), offsetof(scn_surface, c),
         offsetof(scn_universe, n_cells), offsetof(scn_universe, check_overlap), offsetof(scn_universe, rot),
         offsetof(scn_problem, surfaces), offsetof(scn_clerk, kind), offsetof(scn_clerk, mem_addr));
  return 0; }
'''
    tmp = str(tmp_path / "layout_probe")
    with open(tmp + ".c", "w") as f:
        f.write(src)
    subprocess.run(["gcc", "-I", os.path.join(ROOT, "include"), tmp + ".c", "-o", tmp], check=True)
    out = subprocess.run([tmp], capture_output=True, text=True, check=True).stdout.split("\n")
    sizes = [int(x) for x in out[0].split()]
    py = [C.sizeof(T.ParticleState), C.sizeof(T.Box), C.sizeof(T.Universe), C.sizeof(T.MGData),
          C.sizeof(T.Map1D), C.sizeof(T.Clerk), C.sizeof(T.Problem), C.sizeof(T.EigenSettings),
          C.sizeof(T.EigenResult), C.sizeof(engine.Stats)]
    assert sizes == py
    offs = [int(x) for x in out[1].split()]
    assert offs == [T.ParticleState.time.offset, T.ParticleState.broodID.offset, T.Problem.mg.offset]
    fixed = [int(x) for x in out[2].split()]
    assert fixed == [C.sizeof(T.Source), C.sizeof(T.FixedRequest), C.sizeof(T.FixedResult),
                     C.sizeof(T.BatchSettings), T.Source.probG.offset, T.FixedRequest.stop_rel.offset]
    csg = [int(x) for x in out[3].split()]
    assert csg == [C.sizeof(T.Surface), T.Surface.c.offset, T.Universe.n_cells.offset,
                   T.Universe.check_overlap.offset, T.Universe.rot.offset, T.Problem.surfaces.offset,
                   T.Clerk.kind.offset, T.Clerk.mem_addr.offset]


def test_engine_fails_loudly_without_gpu():
    if os.path.exists("/dev/kfd") and os.access("/dev/kfd", os.R_OK):
        pytest.skip("a GPU is visible here")
    with pytest.raises(engine.EngineError):
        engine.device_info()


def test_dictionary_reader_on_reference_format():
    text = """
    type eigenPhysicsPackage;   // comment
    pop 1000; active 5; inactive 2;
    geometry { type geometryStd; boundary (1 1 1 1 1 1);
      surfaces { b { id 1; type box; origin (0.0 0.0 0.0); halfwidth (5.0 5.0 5.0); } }
      universes { root { id 1; type rootUniverse; border 1; fill fuel; } } }
    ! another comment
    """
    d = dictio.parse(text)
    assert d["pop"] == 1000 and d["geometry"]["boundary"] == [1] * 6
    assert d["geometry"]["surfaces"]["b"]["halfwidth"] == [5.0, 5.0, 5.0]
    mg = dictio.read(os.path.join(ROOT, "tests", "golden", "mgMat2"))
    assert mg["numberOfGroups"] == 4 and len(mg["P0"]) == 16 and mg["kappa"] == [202, 193, 180, 200]
    assert dictio.parse(dictio.dumps(d)) == d


def test_configs_compile():
    for name, fn in configs.CONFIGS.items():
        inp, mg = fn(pop=100)
        comp = T.compile_problem(inp, mg)
        assert comp.problem.n_universes >= 1
    inp, mg = configs.pin_lattice_7g(pop=100)
    comp = T.compile_problem(inp, mg)
    assert comp.problem.n_bins == 7 * 2 * 289
    assert comp.bounds == pytest.approx([-10.71] * 3 + [10.71] * 3)


FORTRAN_DRIVER = """
program drive
  use iso_c_binding
  use scone_gpu_mod
  implicit none
  type(c_ptr) :: eng, bat
  integer(c_int) :: rc, rc2
  type(scn_particle_state) :: p
  type(scn_source) :: src
  type(scn_fixed_request) :: req
  type(scn_fixed_result) :: res
  type(scn_batch_settings) :: bs
  rc = scn_create(c_null_ptr, 10_c_int32_t, eng)
  bs % transport = SCN_TRANSPORT_DT
  bs % max_slots = 0
  bs % hbm_budget = 0
  rc2 = scn_batch_create(bs, bat)
  print '(I0,1X,I0,1X,I0,1X,I0,1X,I0,1X,I0,1X,I0)', rc, c_sizeof(p), rc2, c_sizeof(src), c_sizeof(req), &
        c_sizeof(res), c_sizeof(bs)
end program drive
"""


def test_fortran_binding_links_and_calls_the_library(tmp_path):
    """The iso_c_binding module compiles with the image's Fortran compiler (amdflang) and calls
    into libscone_amd.so: scn_create with no problem tables must return an error code."""
    import shutil
    fc = shutil.which("amdflang") or "/opt/rocm/bin/amdflang"
    if not os.path.exists(fc):
        pytest.skip("no Fortran compiler in this image")
    work = str(tmp_path)
    with open(os.path.join(work, "drive.f90"), "w") as f:
        f.write(FORTRAN_DRIVER)
    lib_dir = os.path.join(ROOT, "scone_amd")
    subprocess.run([fc, "-c", os.path.join(ROOT, "include", "scone_gpu_mod.f90"), "-J", work, "-o",
                    os.path.join(work, "mod.o")], check=True, cwd=work)
    subprocess.run([fc, "-I", work, os.path.join(work, "drive.f90"), os.path.join(work, "mod.o"),
                    "-L", lib_dir, "-lscone_amd", f"-Wl,-rpath,{lib_dir}", "-o", os.path.join(work, "drive")],
                   check=True, cwd=work)
    out = subprocess.run([os.path.join(work, "drive")], capture_output=True, text=True, check=True).stdout.split()
    rc, size = int(out[0]), int(out[1])
    assert rc != 0                       # SCN_ERR_NO_DEVICE without a GPU, SCN_ERR_ARG with one
    assert size == C.sizeof(T.ParticleState)
    assert int(out[2]) != 0              # invalid batch settings (or no device)
    assert [int(x) for x in out[3:7]] == [C.sizeof(T.Source), C.sizeof(T.FixedRequest), C.sizeof(T.FixedResult),
                                          C.sizeof(T.BatchSettings)]
