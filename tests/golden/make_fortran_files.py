"""Write the Fortran-record fixtures with a real Fortran compiler.

Emits a small Fortran program of this project (its records follow the write
statements of the reference's wave_to_disk and block_to_disk,
code/space_info.f90:1406-1436 and :1470-1492: sequential unformatted files,
default integer kind, complex(8) matrices), compiles it with the image's
flang and runs it, producing

  wave_hubbard4x2_fortran.dat   psi of the Hubbard 4 x 2 ladder superblock
                                (sys block 3 sites, env block 3 sites, N = 6,
                                S = 0, no truncation): sys_len, env_len of
                                the enlarged blocks, num, dim, then per super
                                sector an empty record, (N, 2S) of both
                                sectors, (sys_dim, env_dim) and the block;
                                element k of the super-basis vector is
                                (1e-3 k, -2e-3 k)
  trun_hubbard4x2_fortran.dat   a truncation of the enlarged sys block of 2
                                sites (len 3, num, dim, up_dif 0, down_dif 0,
                                then per kept sector an empty record,
                                (N, 2S, 0), (row_dim, sdim), the matrix);
                                every other enlarged sector is kept with
                                max(1, dim // 2) columns, element (i, j) of
                                kept sector s is (0.1 s + 0.001 (i + 1),
                                -0.002 (j + 1))
  basis_hubbard4x2_fortran.txt  Get_basis of the sys block of 3 sites and one
                                site in the formatted layout of
                                basis_to_disk (code/space_info.f90:1665-1706:
                                its write statements and edit descriptors
                                (A4,I6), (A11,I4), (A13,I4), (A10,I4,2X,A12,I4)
                                and the empty list-directed record)

The sector tables come from the oracle (test infrastructure).  Run from the
repository root:  python tests/golden/make_fortran_files.py
"""
import os
import shutil
import subprocess
import sys
import tempfile

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from oracle import dmrg, driver  # noqa: E402
from oracle import lattice as lat  # noqa: E402

MODEL = lat.Model(kind="hubbard", Lx=4, Ly=2, U=8.0, t=1.0, flux=0.15, n_elec=6)


def tables():
    run = driver.Run(MODEL, driver.Schedule(kept_max=2000, n_sweeps=0, tol=1e-9))
    run.warm_up()
    sysb, envb = run.sys_blocks[3], run.env_blocks[3]
    sg = dmrg.StepGeometry(MODEL, 3, 3)
    se = dmrg.enlarge_in_geometry(sysb, MODEL, "sys", sg)
    ee = dmrg.enlarge_in_geometry(envb, MODEL, "env", sg)
    sb = dmrg.super_basis(se.basis, ee.basis, MODEL.n_elec)
    wave = []
    off = 0
    for (x, y) in sb.pairs:
        r, c = se.basis.dims[x], ee.basis.dims[y]
        wave.append((se.basis.sectors[x], ee.basis.sectors[y], r, c, off))
        off += r * c
    eb = dmrg.enlarge_basis(run.sys_blocks[2].sectors, run.sys_blocks[2].dims, dmrg.site_sectors("hubbard"))
    trun = []
    for a, (q, d) in enumerate(zip(eb.sectors, eb.dims)):
        if a % 2 == 0:
            trun.append((q, d, max(1, d // 2)))
    basis = dmrg.enlarge_basis(sysb.sectors, sysb.dims, dmrg.site_sectors("hubbard"))
    return wave, off, trun, basis


def basis_source(basis, length):
    """The write statements of basis_to_disk for one enlarged basis."""
    L = ["  open(13, file='basis_hubbard4x2_fortran.txt', status='replace')",
         "  write(13,101) \"len=\", %d" % length, "  write(13,101) \"num=\", %d" % len(basis.sectors),
         "  write(13,101) \"dim=\", %d" % sum(basis.dims)]
    for (N, S2), d, subs in zip(basis.sectors, basis.dims, basis.subs):
        L += ["  write(13,*)", "  write(13,101) \"idx=\", %d" % len(subs), "  write(13,101) \"num=\", %d" % len(subs),
              "  write(13,101) \"dim=\", %d" % d, "  write(13,102) \"new_num_up=\", %d" % N,
              "  write(13,103) \"new_num_down=\", %d" % S2]
        for (b, st, pos, sdim) in subs:
            bN, bS = basis.bl_sectors[b]
            sN, sS = basis.st_sectors[st]
            L += ["  write(13,101) \"pos=\", %d" % pos, "  write(13,101) \"dim=\", %d" % sdim,
                  "  write(13,104) \"bl_num_up=\", %d, \"bl_num_down=\", %d" % (bN, bS),
                  "  write(13,104) \"st_num_up=\", %d, \"st_num_down=\", %d" % (sN, sS)]
    L += ["  close(13)"]
    return L


def fortran_source(wave, dim, trun, basis):
    L = ["program write_fixtures", "  implicit none", "  integer :: i, j, k", "  complex(8), allocatable :: v(:, :)",
         "  open(11, file='wave_hubbard4x2_fortran.dat', form='unformatted', status='replace')",
         "  write(11) 4", "  write(11) 4", "  write(11) %d" % len(wave), "  write(11) %d" % dim]
    for (sq, eq, r, c, off) in wave:
        L += ["  write(11)", "  write(11) %d, %d" % sq, "  write(11) %d, %d" % eq, "  write(11) %d, %d" % (r, c),
              "  allocate(v(%d, %d))" % (r, c),
              "  do j = 1, %d" % c, "    do i = 1, %d" % r,
              "      k = %d + (i - 1) + (j - 1) * %d" % (off, r),
              "      v(i, j) = cmplx(1.0d-3 * k, -2.0d-3 * k, kind=8)", "    end do", "  end do",
              "  write(11) v", "  deallocate(v)"]
    L += ["  close(11)",
          "  open(12, file='trun_hubbard4x2_fortran.dat', form='unformatted', status='replace')",
          "  write(12) 3", "  write(12) %d" % len(trun), "  write(12) %d" % sum(k for _, _, k in trun),
          "  write(12) 0", "  write(12) 0"]
    for s, (q, d, keep) in enumerate(trun):
        L += ["  write(12)", "  write(12) %d, %d, 0" % q, "  write(12) %d, %d" % (d, keep),
              "  allocate(v(%d, %d))" % (d, keep),
              "  do j = 1, %d" % keep, "    do i = 1, %d" % d,
              "      v(i, j) = cmplx(0.1d0 * %d + 1.0d-3 * i, -2.0d-3 * j, kind=8)" % s, "    end do", "  end do",
              "  write(12) v", "  deallocate(v)"]
    L += ["  close(12)"]
    L += basis_source(basis, 4)
    L += ["101 format(A4,I6)", "102 format(A11,I4)", "103 format(A13,I4)", "104 format(A10,I4,2X,A12,I4)",
          "end program write_fixtures"]
    return "\n".join(L) + "\n"


def main():
    flang = shutil.which("flang") or shutil.which("amdflang") or "/opt/rocm/bin/amdflang"
    wave, dim, trun, basis = tables()
    with tempfile.TemporaryDirectory() as tmp:
        src = os.path.join(tmp, "w.f90")
        open(src, "w").write(fortran_source(wave, dim, trun, basis))
        exe = os.path.join(tmp, "w")
        subprocess.run([flang, "-o", exe, src], check=True)
        subprocess.run([exe], check=True, cwd=HERE)
    print("wrote", *(os.path.join(HERE, f) for f in ("wave_hubbard4x2_fortran.dat", "trun_hubbard4x2_fortran.dat",
                                                      "basis_hubbard4x2_fortran.txt")))


if __name__ == "__main__":
    main()
