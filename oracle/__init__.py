"""CPU oracle for the SU(2) DMRG inner loop -- TEST INFRASTRUCTURE ONLY.

This package is a plain numpy restatement of the reference algorithm
(cfengno1/Hofstadter_Hubbard_TSC, Fortran).  It exists so that tests, the
smoke check and the CPU-baseline leg of the benchmark can check the HIP
product path against an independent implementation.  Nothing in the product
(the ``hofstadter_hubbard_tsc_amd`` package or ``libhhdmrg.so``) imports,
links or calls it.

Pins (how each part of the oracle is checked, see ``tests/test_oracle_*.py``):

* ``su2.w6j``            -- sympy ``wigner_6j`` on every argument set used,
                            closed forms with a zero argument, orthogonality.
* ``lattice``            -- bond counts / coordination on square and
                            triangular lattices, Landau-gauge plaquette flux
                            (product of phases around each plaquette).
* ``dmrg`` matvec        -- the spectrum of the untruncated reduced
                            superblock Hamiltonian equals the S=0 spectrum of
                            an independent Jordan-Wigner exact diagonalisation
                            (``ed.py``) on Hubbard / t-J lattices with flux;
                            Hermiticity of H.
* ``dmrg`` growth,       -- with no truncation every DMRG step reproduces the
  renormalisation,          exact ground-state energy of its geometry (ED);
  wave transformation       the wave transformation is an isometry and
                            <psi'|H'|psi'> = <psi|H|psi> for an untruncated
                            rotation.
* Lanczos                -- numpy ``eigh`` on explicit Hermitian matrices;
                            stopping rule tested as a specification.
* ``lanczos_levels``     -- the lowest eigenvalues of ``eigh`` and the S=0
                            levels of the ED on an untruncated superblock;
                            the reference's real-only projection, planted as
                            a mistake, returns E0 again for level 2.
* mixed density matrix   -- (0.1 rho_1 + 1.9 rho_2)/2 block by block with the
                            single-precision weights, trace 1 - 1.1e-8.
* density / truncation   -- trace 1, numpy ``eigh``; kept-count rule on
                            synthetic spectra (ties, kept_min/kept_max edges).

Parity unpinned: the finite-m sweep trajectory of the reference itself
(per-step energies, kept counts) -- the reference Fortran needs Intel MKL, an
external library, so it is not built here.
"""
