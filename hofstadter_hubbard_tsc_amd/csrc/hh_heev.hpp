// Hand-written eigensolver of the reduced density matrix, all sectors of one
// truncation at once (stands in for the per-sector ZHEEV of
// Get_truncation_operator, code/get_den.f90:242).
//
//   values():  every sector's Hermitian block is reduced to a real symmetric
//              tridiagonal T = Q^H A Q by Householder reflectors (one launch
//              per column for all sectors together, the reflectors kept in
//              the block's strictly lower part), then every eigenvalue of
//              every T by bisection on Sturm counts (one thread each);
//   vectors(): for the kept eigenvalues only (the top keep[i] of sector i):
//              inverse iteration on T - lambda I (one thread per cluster of
//              eigenvalues closer than 1e-12 of the sector's scale, random
//              starts, Gram-Schmidt inside a cluster; kept eigenvalues in
//              the numerical null space above the sector's smallest get an
//              orthonormal basis of the complement of the others), Newton-Schulz
//              orthonormalisation of each sector's kept vectors
//              (Z <- Z (3I - Z^T Z) / 2 until |Z^T Z - I| <= 1e-12), and
//              the back-transformation U = Q Z in blocks of 32 reflectors
//              (U -= V T V^H U)
//              (one wave per vector, all reflectors in one launch).  U is
//              written as the last keep[i] columns of the sector's block in
//              ascending eigenvalue order (the layout ZHEEV leaves), every
//              other entry of the block zero.
// Every reduction runs in a fixed order: results are bit-reproducible.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>
#include <vector>

namespace hh {

// Largest sector taken (the Householder step keeps one column of the sector
// in LDS); a truncation with a larger sector uses the library eigensolver.
constexpr int HEEV_MAXN = 3072;

// Sector record shared by the host and the kernels.
struct HeevSec {
  int64_t a;      // offset of the n x n block (complex elements, column-major)
  int64_t c;      // offset of the sector's n-vectors (d, e, tau, v, p, eigenvalues)
  int64_t z;      // offset of its n x keep real blocks (Z, Z2) and U (n x keep complex)
  int64_t w;      // offset of its inverse-iteration scratch (5 n keep doubles)
  int64_t g;      // offset of its keep x keep Gram matrix
  int64_t cb;     // offset of the T factors of its blocks of reflectors (NB x NB each)
  int64_t wb;     // offset of its NB x keep product V^H U of the back-transformation
  int32_t n, ch0, nch, keep;
};

// Workgroup record of a Householder step: the sector's block and vector
// offsets, size and share range, and the workgroup's column chunk q.
struct HeevWg {
  int64_t a, c;
  int32_t n, ch0, nch, q;
};

// grow-only device buffer
struct HeevBuf {
  void* p = nullptr;
  size_t bytes = 0;
  void ensure(size_t b);
  ~HeevBuf();
  template <typename T>
  T* as() const { return (T*)p; }
};

struct HeevWork {
  // statistics of the last call
  int ns_iterations = 0;        // Newton-Schulz steps taken
  int max_cluster = 0;          // largest cluster of kept eigenvalues
  int64_t unresolved = 0;       // kept vectors completed in a numerical null space
  std::string fail_info;        // why the last vectors() call returned false
  // with timing set, each phase is synchronised and timed (ms): Householder
  // reduction, bisection, inverse iteration, orthonormalisation,
  // back-transformation
  bool timing = false;
  double ms[5] = {0, 0, 0, 0, 0};
  double ortho_error = 0.0;     // max |Z^T Z - I| at the end

  // Eigenvalues (ascending per sector, concatenated in sector order) of the
  // Hermitian blocks at A + off[i] (dim[i] x dim[i], column-major; the upper
  // triangle is read).  A is overwritten with the reflectors.
  void values(hipStream_t s, double2* A, const std::vector<int64_t>& off, const std::vector<int>& dim,
              double* evals_dev);
  // Eigenvectors of the top keep[i] eigenvalues of each sector of the last
  // values() call, written into A as described above.  Returns false when
  // the kept vectors could not be orthonormalised (A is then undefined).
  bool vectors(hipStream_t s, double2* A, const std::vector<int>& keep, const double* evals_dev);

 private:
  std::vector<HeevSec> sec_;
  std::vector<int64_t> off_;
  std::vector<int> dim_;
  int64_t ncol_ = 0, nblk_ = 0;
  int nmax_ = 0, nwg_ = 0, ntile_ = 0;
  // host staging of the tables (alive until the copies are done)
  std::vector<HeevWg> h_wgr_;
  std::vector<int32_t> h_tile_, h_owner_, h_vown_, h_clus_, h_gtile_, h_ztile_, h_bt_, h_nu_, h_ftile_,
      h_xtile_, h_ucol_, h_rg_, h_wyw_, h_wyu_;
  std::vector<int64_t> h_zoff_;
  HeevBuf secs_, wgs_, tiles_, owner_, d_, e_, tau_, vb_, pb_, part_;
  HeevBuf vown_, scratch_, Z_, Z2_, G_, gtile_, gerr_, ztile_, U_, bt_, scl_, nu_, ftile_, xtile_, ucol_, rg_, cg_,
      wyu_, zoff_, Wb_;
  int npart_ = 0;
};

}  // namespace hh
