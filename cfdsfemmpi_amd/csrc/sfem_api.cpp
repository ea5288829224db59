// sfem_api.cpp -- host side of the C ABI (include/sfem.h).
//
// Owns the context, builds the DOF numbering, the CSR pattern, the
// element-to-nonzero map and the element colouring on the host (multi-threaded),
// and drives the HIP kernels for the state updates, the element loops and the
// Krylov solves.  See each ABI function's comment in sfem.h for the reference
// interface it replaces.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <thread>
#include <vector>

#include "sfem_internal.hpp"
#include "sfem_linalg.hpp"
#include "sfem_state.hpp"

namespace sfem {

static thread_local std::string g_err;

int set_error(int code, const char* msg) {
  g_err = msg ? msg : "";
  return code;
}

int set_hip_error(hipError_t e, const char* what) {
  g_err = std::string(what) + ": " + hipGetErrorString(e);
  return SFEM_ERR_HIP;
}

#define HIP_OK(x)                                     \
  do {                                                \
    hipError_t _e = (x);                              \
    if (_e != hipSuccess) return set_hip_error(_e, #x); \
  } while (0)

// Time-integration coefficients of SetTimeParametersFluid (UtilitiesGeneral.cpp:5-209).
static int time_params(int tis, double rho, double dt, double* td) {
  for (int i = 0; i < 16; ++i) td[i] = 0.0;
  td[0] = dt;
  double af, am, g;
  switch (tis) {
    case 0:
      td[1] = td[2] = td[3] = td[4] = 1.0;
      return 0;
    case 1:
      af = 1.0 / (1.0 + rho); am = 1.0; g = 1.0;
      td[1] = am; td[2] = af; td[3] = am; td[4] = g;
      td[5] = af * dt; td[6] = am / g; td[7] = 1.0 - am / g; td[8] = 0.0;
      td[9] = 1.0 / dt; td[10] = -td[9];
      return 0;
    case 2:
    case 4:
      af = 1.0 / (1.0 + rho);
      am = 0.5 * (3.0 - rho) / (1.0 + rho);
      g = 0.5 + am - af;
      td[1] = am; td[2] = af; td[3] = am; td[4] = g;
      td[5] = af * dt; td[6] = am / g; td[7] = 1.0 - am / g; td[8] = am / g / dt;
      td[9] = 1.0 / g / dt; td[10] = -td[9]; td[15] = 1.0 - 1.0 / g;
      return 0;
    case 3:
    case 5:
    case 6:
    case 7:
      td[1] = td[2] = td[3] = td[4] = 1.0;
      td[5] = dt; td[6] = 1.0; td[7] = 0.0; td[8] = 1.0 / dt;
      if (tis == 3) { td[9] = 1.0 / dt; td[10] = -1.0 / dt; }
      if (tis == 5) { td[9] = 1.5 / dt; td[10] = -2.0 / dt; td[11] = 0.5 / dt; }
      if (tis == 6) {
        td[9] = (11.0 / 6.0) / dt; td[10] = -3.0 / dt; td[11] = 1.5 / dt; td[12] = -(1.0 / 3.0) / dt;
      }
      if (tis == 7) {
        td[9] = (25.0 / 12.0) / dt; td[10] = -4.0 / dt; td[11] = 3.0 / dt;
        td[12] = -(4.0 / 3.0) / dt; td[13] = (1.0 / 4.0) / dt;
      }
      return 0;
    default:
      return -1;
  }
}

template <typename T>
static int dalloc(T** p, size_t n) {
  *p = nullptr;
  hipError_t e = hipMalloc((void**)p, std::max<size_t>(n, 1) * sizeof(T));
  if (e != hipSuccess) return set_hip_error(e, "hipMalloc");
  return SFEM_OK;
}

template <typename T>
static void dfree(T*& p) {
  if (p) (void)hipFree((void*)p);
  p = nullptr;
}

static int64_t nstate(const sfem_ctx* c) { return c->nNode * c->dims.ndof; }

static AsmView asm_view(sfem_ctx* c) {
  AsmView v;
  v.coords = c->d_coords;
  v.conn = c->d_conn;
  v.charlen = c->d_charlen;
  v.solnPrev = c->d_prev;
  v.solnCur = c->d_cur;
  v.solnDotCur = c->d_dotCur;
  v.solnApplied = c->d_applied;
  v.nodeInfo = c->d_nodeInfo;
  v.colofs = c->d_colofs;
  v.sconn = c->d_sconn;
  v.vals = c->d_vals;
  v.vals32 = c->d_vals32;
  v.blkofs = c->d_blkofs;
  v.kb = c->d_kb;
  v.kb32 = c->d_kb32;
  v.rhs = c->d_rhs;
  v.probeK = c->d_probeK;
  v.probeF = c->d_probeF;
  v.elem_mu = c->d_elem_mu;
  return v;
}

static ElemParams elem_params(const sfem_ctx* c) {
  ElemParams p;
  p.rho = c->elemData[0];
  p.mu = c->elemData[1];
  p.af = c->td[2];
  p.acceFact = c->td[8];
  p.stab[0] = c->elemData[8];
  p.stab[1] = c->elemData[9];
  p.stab[2] = c->elemData[10];
  return p;
}

static TimeCoeffs time_coeffs(const sfem_ctx* c) {
  return TimeCoeffs{c->td[1], c->td[2], c->td[9], c->td[10], c->td[11], c->td[12], c->td[13], c->td[15]};
}

enum { ST_PREP = 1, ST_PATTERN = 2, ST_PARAMS = 4, ST_ASSEMBLED = 8, ST_DBC = 16 };

}  // namespace sfem

using namespace sfem;

extern "C" {

const char* sfem_version(void) { return "sfem-mi355x 0.1 (gfx950)"; }

const char* sfem_last_error(void) { return g_err.c_str(); }

sfem_ctx* sfem_create(int ndim, int npElem, int64_t nNode, int64_t nElem, const double* coords,
                      const int32_t* conn, int device) {
  int et = -1;
  if (ndim == 2 && npElem == 3) et = TRIA3;
  else if (ndim == 2 && npElem == 4) et = QUAD4;
  else if (ndim == 3 && npElem == 4) et = TETRA4;
  else if (ndim == 3 && npElem == 8) et = HEXA8;
  if (et < 0) { set_error(SFEM_ERR_UNSUPPORTED, "unsupported (ndim, npElem)"); return nullptr; }
  if (nNode <= 0 || nElem <= 0 || !coords || !conn) { set_error(SFEM_ERR_ARG, "empty mesh"); return nullptr; }
  if (nNode * (ndim + 1) >= (int64_t)INT32_MAX) { set_error(SFEM_ERR_ARG, "mesh too large for int32 DOF ids"); return nullptr; }
  for (int64_t i = 0; i < nElem * npElem; ++i)
    if (conn[i] < 0 || conn[i] >= nNode) { set_error(SFEM_ERR_ARG, "connectivity out of range"); return nullptr; }
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) { set_error(SFEM_ERR_HIP, "no HIP device"); return nullptr; }
  sfem_ctx* c = new sfem_ctx;
  if (device >= 0) {
    if (device >= ndev) { delete c; set_error(SFEM_ERR_ARG, "device index out of range"); return nullptr; }
    c->device = device;
    (void)hipSetDevice(device);
  } else {
    (void)hipGetDevice(&c->device);
  }
  // the allocation-free entry point's scratch array: this device's address of it, looked
  // up once per device with the device current
  if (la_scratch_init(c->device) != SFEM_OK) {
    std::string msg = g_err;
    delete c;
    set_error(SFEM_ERR_HIP, msg.c_str());
    return nullptr;
  }
  c->et = et;
  c->dims = elem_dims(et);
  c->nNode = nNode;
  c->nElem = nElem;
  c->coords.assign(coords, coords + nNode * ndim);
  c->conn.assign(conn, conn + nElem * npElem);
  bool ok = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) == hipSuccess;
  ok = ok && hipEventCreate(&c->ev0) == hipSuccess && hipEventCreate(&c->ev1) == hipSuccess &&
       hipEventCreate(&c->ev2) == hipSuccess && hipEventCreate(&c->ev3) == hipSuccess;
  ok = ok && dalloc(&c->d_coords, c->coords.size()) == SFEM_OK && dalloc(&c->d_conn, c->conn.size()) == SFEM_OK &&
       dalloc(&c->d_charlen, nElem) == SFEM_OK && dalloc(&c->d_vol, nElem) == SFEM_OK &&
       dalloc(&c->d_bad, 1) == SFEM_OK;
  const int64_t ns = nNode * c->dims.ndof;
  double** sv[] = {&c->d_soln, &c->d_prev, &c->d_prev2, &c->d_prev3, &c->d_prev4, &c->d_dot,
                   &c->d_dotPrev, &c->d_cur, &c->d_dotCur, &c->d_applied};
  for (double** q : sv) ok = ok && dalloc(q, ns) == SFEM_OK && hipMemset(*q, 0, ns * sizeof(double)) == hipSuccess;
  const int nsz = c->dims.npe * c->dims.ndof;
  ok = ok && dalloc(&c->d_probeK, nsz * nsz) == SFEM_OK && dalloc(&c->d_probeF, nsz) == SFEM_OK;
  ok = ok && hipHostMalloc((void**)&c->h_pinned, 64 * sizeof(double), hipHostMallocDefault) == hipSuccess;
  if (ok) {
    ok = hipMemcpy(c->d_coords, coords, c->coords.size() * sizeof(double), hipMemcpyHostToDevice) == hipSuccess &&
         hipMemcpy(c->d_conn, conn, c->conn.size() * sizeof(int32_t), hipMemcpyHostToDevice) == hipSuccess;
  }
  if (!ok) {
    std::string msg = g_err.empty() ? "device allocation failed" : g_err;
    sfem_destroy(c);
    set_error(SFEM_ERR_NOMEM, msg.c_str());
    return nullptr;
  }
  return c;
}

void sfem_destroy(sfem_ctx* c) {
  if (!c) return;
  (void)hipSetDevice(c->device);
  if (c->stream) (void)hipStreamSynchronize(c->stream);
  if (c->la) la_destroy(c);
  if (c->batch) batch_destroy(c);
  dfree(c->d_elem_mu);
  dfree(c->d_coords); dfree(c->d_conn); dfree(c->d_charlen); dfree(c->d_vol); dfree(c->d_bad);
  dfree(c->d_nodeInfo); dfree(c->d_colofs); dfree(c->d_sconn); dfree(c->d_elemByColor); dfree(c->d_rowptr32);
  dfree(c->d_colidx); dfree(c->d_assy); dfree(c->d_dbc_ind); dfree(c->d_dbc_val); dfree(c->d_dbc_mult);
  dfree(c->d_vals); dfree(c->d_vals32); dfree(c->d_rhs); dfree(c->d_du);
  dfree(c->d_blkofs); dfree(c->d_natofs); dfree(c->d_blkPtr); dfree(c->d_kb); dfree(c->d_kb32);
  dfree(c->d_soln); dfree(c->d_prev); dfree(c->d_prev2); dfree(c->d_prev3); dfree(c->d_prev4);
  dfree(c->d_dot); dfree(c->d_dotPrev); dfree(c->d_cur); dfree(c->d_dotCur); dfree(c->d_applied);
  dfree(c->d_probeK); dfree(c->d_probeF); dfree(c->d_cand);
  if (c->h_pinned) (void)hipHostFree(c->h_pinned);
  hipEvent_t evs[] = {c->ev0, c->ev1, c->ev2, c->ev3};
  for (hipEvent_t e : evs)
    if (e) (void)hipEventDestroy(e);
  if (c->stream) (void)hipStreamDestroy(c->stream);
  delete c;
}

int sfem_set_dirichlet(sfem_ctx* c, int64_t nDBC, const int32_t* node, const int32_t* dof,
                       const double* value) {
  if (!c) return set_error(SFEM_ERR_ARG, "null context");
  if (c->status & ST_PATTERN) return set_error(SFEM_ERR_STATE, "Dirichlet data must be set before the pattern");
  if (nDBC < 0 || (nDBC > 0 && (!node || !dof || !value))) return set_error(SFEM_ERR_ARG, "bad Dirichlet arrays");
  const int ndof = c->dims.ndof;
  c->dbc_ind_raw.clear();
  c->dbc_val_raw.clear();
  std::vector<int32_t> slot(c->nNode * ndof, -1);
  c->dbc_ind.clear(); c->dbc_val.clear(); c->dbc_mult.clear();
  for (int64_t k = 0; k < nDBC; ++k) {
    if (node[k] < 0 || node[k] >= c->nNode || dof[k] < 0 || dof[k] >= ndof)
      return set_error(SFEM_ERR_ARG, "Dirichlet entry out of range");
    const int32_t ind = node[k] * ndof + dof[k];
    c->dbc_ind_raw.push_back(ind);
    c->dbc_val_raw.push_back(value[k]);
    if (slot[ind] < 0) {
      slot[ind] = (int32_t)c->dbc_ind.size();
      c->dbc_ind.push_back(ind);
      c->dbc_val.push_back(value[k]);
      c->dbc_mult.push_back(1);
    } else {
      c->dbc_val[slot[ind]] = value[k];  // the last listing's value wins (StabFEM.cpp:398-403)
      c->dbc_mult[slot[ind]] += 1;
    }
  }
  (void)hipSetDevice(c->device);
  dfree(c->d_dbc_ind); dfree(c->d_dbc_val); dfree(c->d_dbc_mult);
  const size_t nu = c->dbc_ind.size();
  int rc;
  if ((rc = dalloc(&c->d_dbc_ind, nu)) || (rc = dalloc(&c->d_dbc_val, nu)) || (rc = dalloc(&c->d_dbc_mult, nu)))
    return rc;
  if (nu) {
    HIP_OK(hipMemcpy(c->d_dbc_ind, c->dbc_ind.data(), nu * sizeof(int32_t), hipMemcpyHostToDevice));
    HIP_OK(hipMemcpy(c->d_dbc_val, c->dbc_val.data(), nu * sizeof(double), hipMemcpyHostToDevice));
    HIP_OK(hipMemcpy(c->d_dbc_mult, c->dbc_mult.data(), nu * sizeof(int32_t), hipMemcpyHostToDevice));
  }
  c->status |= ST_DBC;
  return SFEM_OK;
}

int sfem_prepare_elements(sfem_ctx* c) {
  if (!c) return set_error(SFEM_ERR_ARG, "null context");
  HIP_OK(hipSetDevice(c->device));
  const int64_t big = INT64_MAX;
  HIP_OK(hipMemcpyAsync(c->d_bad, &big, sizeof(int64_t), hipMemcpyHostToDevice, c->stream));
  HIP_OK(launch_prepare(c->et, c->stream, c->nElem, c->d_coords, c->d_conn, c->d_charlen, c->d_vol, c->d_bad));
  int64_t bad = 0;
  HIP_OK(hipMemcpyAsync(&bad, c->d_bad, sizeof(int64_t), hipMemcpyDeviceToHost, c->stream));
  HIP_OK(hipStreamSynchronize(c->stream));
  if (bad != INT64_MAX) {
    char msg[128];
    std::snprintf(msg, sizeof msg, "negative Jacobian in element %lld", (long long)bad);
    return set_error(SFEM_ERR_NEGJAC, msg);
  }
  c->status |= ST_PREP;
  return SFEM_OK;
}

int sfem_build_pattern(sfem_ctx* c, int64_t* nfree_out, int64_t* nnz_out) {
  if (!c) return set_error(SFEM_ERR_ARG, "null context");
  HIP_OK(hipSetDevice(c->device));
  const int ndof = c->dims.ndof, npe = c->dims.npe;
  const int64_t nNode = c->nNode, nElem = c->nElem;
  // DOF numbering: node-major, prescribed DOFs skipped (StabFEM2.cpp:65-81)
  std::vector<char> fixed(nNode * ndof, 0);
  for (int32_t ind : c->dbc_ind) fixed[ind] = 1;
  c->nodeDof.assign(nNode * ndof, -1);
  c->assyForSoln.clear();
  int64_t nf = 0;
  std::vector<int32_t> mask(nNode, 0), cnt(nNode, 0), firstRow(nNode, -1);
  for (int64_t i = 0; i < nNode; ++i)
    for (int d = 0; d < ndof; ++d)
      if (!fixed[i * ndof + d]) {
        if (firstRow[i] < 0) firstRow[i] = (int32_t)nf;
        c->nodeDof[i * ndof + d] = (int32_t)nf;
        c->assyForSoln.push_back((int32_t)(i * ndof + d));
        mask[i] |= 1 << d;
        cnt[i]++;
        ++nf;
      }
  c->nfree = nf;
  // node -> elements
  std::vector<int64_t> nePtr(nNode + 1, 0);
  for (int64_t q = 0; q < nElem * npe; ++q) nePtr[c->conn[q] + 1]++;
  for (int64_t i = 0; i < nNode; ++i) nePtr[i + 1] += nePtr[i];
  std::vector<int32_t> neList(nElem * npe);
  {
    std::vector<int64_t> pos(nePtr.begin(), nePtr.end() - 1);
    for (int64_t e = 0; e < nElem; ++e)
      for (int a = 0; a < npe; ++a) neList[pos[c->conn[e * npe + a]]++] = (int32_t)e;
  }
  // node adjacency restricted to nodes with free DOFs; the row pattern of node i is the
  // union of the free DOFs of its neighbours in increasing node order (= increasing DOF id,
  // as the per-row std::set of StabFEM2.cpp:192-227 orders them)
  std::vector<std::vector<int32_t>> nbr(nNode);
  std::vector<int32_t> rowlen(nNode, 0);
  parallel_for(nNode, [&](int64_t lo, int64_t hi) {
    std::vector<int32_t> tmp;
    for (int64_t i = lo; i < hi; ++i) {
      if (!cnt[i]) continue;
      tmp.clear();
      for (int64_t q = nePtr[i]; q < nePtr[i + 1]; ++q) {
        const int64_t e = neList[q];
        for (int a = 0; a < npe; ++a) {
          const int32_t j = c->conn[e * npe + a];
          if (cnt[j]) tmp.push_back(j);
        }
      }
      std::sort(tmp.begin(), tmp.end());
      tmp.erase(std::unique(tmp.begin(), tmp.end()), tmp.end());
      int32_t len = 0;
      for (int32_t j : tmp) len += cnt[j];
      rowlen[i] = len;
      nbr[i] = tmp;
    }
  });
  c->rowptr.assign(nf + 1, 0);
  std::vector<int64_t> rowStart(nNode, 0);
  {
    int64_t r = 0;
    for (int64_t i = 0; i < nNode; ++i) {
      rowStart[i] = c->rowptr[r];
      for (int k = 0; k < cnt[i]; ++k, ++r) c->rowptr[r + 1] = c->rowptr[r] + rowlen[i];
    }
  }
  c->nnz = c->rowptr[nf];
  if (c->nnz >= (int64_t)INT32_MAX) return set_error(SFEM_ERR_UNSUPPORTED, "nnz exceeds int32 range");
  c->colidx.resize(c->nnz);
  parallel_for(nNode, [&](int64_t lo, int64_t hi) {
    for (int64_t i = lo; i < hi; ++i) {
      if (!cnt[i]) continue;
      int64_t q = rowStart[i];
      for (int32_t j : nbr[i])
        for (int d = 0; d < ndof; ++d)
          if ((mask[j] >> d) & 1) c->colidx[q++] = c->nodeDof[j * ndof + d];
      for (int k = 1; k < cnt[i]; ++k)
        std::copy(c->colidx.begin() + rowStart[i], c->colidx.begin() + rowStart[i] + rowlen[i],
                  c->colidx.begin() + rowStart[i] + (int64_t)k * rowlen[i]);
    }
  });
  // element -> nonzero map per node pair: offset of node j's first free column in node i's rows
  std::vector<int32_t> colofs(nElem * npe * npe, -1);
  parallel_for(nElem, [&](int64_t lo, int64_t hi) {
    for (int64_t e = lo; e < hi; ++e)
      for (int a = 0; a < npe; ++a) {
        const int32_t i = c->conn[e * npe + a];
        if (!cnt[i]) continue;
        const auto& nb = nbr[i];
        for (int b = 0; b < npe; ++b) {
          const int32_t j = c->conn[e * npe + b];
          if (!cnt[j]) continue;
          int32_t off = 0;
          for (int32_t x : nb) {
            if (x == j) break;
            off += cnt[x];
          }
          colofs[(e * npe + a) * npe + b] = off;
        }
      }
  });
  // node-pair blocks: node block i (nodes with free dofs, node order) owns the pairs
  // blkPtr[i] .. blkPtr[i + 1], one per neighbour node in increasing node order
  std::vector<int64_t> nodePair(nNode + 1, 0);
  for (int64_t i = 0; i < nNode; ++i) nodePair[i + 1] = nodePair[i] + (cnt[i] ? (int64_t)nbr[i].size() : 0);
  c->nbpairs = nodePair[nNode];
  if (c->nbpairs >= ((int64_t)1 << 30) / 16)
    return set_error(SFEM_ERR_UNSUPPORTED, "node-pair blocks exceed the element map's range");
  c->blkPtr.clear();
  for (int64_t i = 0; i < nNode; ++i)
    if (cnt[i]) c->blkPtr.push_back(nodePair[i]);
  c->blkPtr.push_back(c->nbpairs);
  std::vector<int32_t> natofs(c->nbpairs);
  for (int64_t i = 0; i < nNode; ++i) {
    if (!cnt[i]) continue;
    int32_t off = 0;
    for (size_t t = 0; t < nbr[i].size(); ++t) {
      const int32_t j = nbr[i][t];
      natofs[nodePair[i] + t] = off | (cnt[j] << 24);
      off += cnt[j];
    }
  }
  std::vector<int32_t> blkofs(nElem * npe * npe, -1);
  parallel_for(nElem, [&](int64_t lo, int64_t hi) {
    for (int64_t e = lo; e < hi; ++e)
      for (int a = 0; a < npe; ++a) {
        const int32_t i = c->conn[e * npe + a];
        if (!cnt[i]) continue;
        const auto& nb = nbr[i];
        for (int b = 0; b < npe; ++b) {
          const int32_t j = c->conn[e * npe + b];
          if (!cnt[j]) continue;
          const int64_t t = std::lower_bound(nb.begin(), nb.end(), j) - nb.begin();
          blkofs[(e * npe + a) * npe + b] = (int32_t)(nodePair[i] + t);
        }
      }
  });
  // greedy element colouring: no two elements of a colour share a node
  std::vector<uint64_t> used(nNode, 0);
  std::vector<int32_t> color(nElem);
  int ncol = 0;
  for (int64_t e = 0; e < nElem; ++e) {
    uint64_t u = 0;
    for (int a = 0; a < npe; ++a) u |= used[c->conn[e * npe + a]];
    if (u == ~0ull) return set_error(SFEM_ERR_UNSUPPORTED, "element colouring needs more than 64 colours");
    int col = 0;
    while ((u >> col) & 1ull) ++col;
    color[e] = col;
    ncol = std::max(ncol, col + 1);
    for (int a = 0; a < npe; ++a) used[c->conn[e * npe + a]] |= 1ull << col;
  }
  c->ncolors = ncol;
  c->colorPtr.assign(ncol + 1, 0);
  for (int64_t e = 0; e < nElem; ++e) c->colorPtr[color[e] + 1]++;
  for (int k = 0; k < ncol; ++k) c->colorPtr[k + 1] += c->colorPtr[k];
  std::vector<int32_t> byColor(nElem);
  {
    std::vector<int32_t> pos(c->colorPtr.begin(), c->colorPtr.end() - 1);
    for (int64_t e = 0; e < nElem; ++e) byColor[pos[color[e]]++] = (int32_t)e;
  }
  // the element loop walks colour classes slot by slot: the element -> nonzero map and
  // the connectivity are stored in that order, so each slot's entries are contiguous
  std::vector<int32_t> scolofs((size_t)nElem * npe * npe), sconn((size_t)nElem * npe);
  std::vector<int32_t> sblkofs((size_t)nElem * npe * npe);
  parallel_for(nElem, [&](int64_t lo, int64_t hi) {
    for (int64_t sl = lo; sl < hi; ++sl) {
      const int64_t e = byColor[sl];
      std::copy(colofs.begin() + e * npe * npe, colofs.begin() + (e + 1) * npe * npe,
                scolofs.begin() + sl * npe * npe);
      std::copy(blkofs.begin() + e * npe * npe, blkofs.begin() + (e + 1) * npe * npe,
                sblkofs.begin() + sl * npe * npe);
      std::copy(c->conn.begin() + e * npe, c->conn.begin() + (e + 1) * npe, sconn.begin() + sl * npe);
    }
  });
  std::vector<int32_t>().swap(colofs);
  std::vector<int32_t>().swap(blkofs);
  // the first element (in colour-slot order, the order the colour launches run in) to
  // reach a block stores it (bit 30) instead of adding to it, so the block form needs no
  // clearing; every block is some element's node pair
  {
    std::vector<uint8_t> seen(c->nbpairs, 0);
    for (size_t k = 0; k < sblkofs.size(); ++k) {
      const int32_t q = sblkofs[k];
      if (q >= 0 && !seen[q]) {
        seen[q] = 1;
        sblkofs[k] = q | (1 << 30);
      }
    }
    for (int64_t q = 0; q < c->nbpairs; ++q)
      if (!seen[q]) return set_error(SFEM_ERR_STATE, "a node-pair block no element reaches");
  }
  // node info {rowStart, rowlen, mask, firstRow}
  std::vector<int32_t> info(nNode * 4);
  for (int64_t i = 0; i < nNode; ++i) {
    info[4 * i] = (int32_t)rowStart[i];
    info[4 * i + 1] = rowlen[i];
    info[4 * i + 2] = mask[i];
    info[4 * i + 3] = firstRow[i];
  }
  std::vector<int32_t> rp32(nf + 1);
  for (int64_t r = 0; r <= nf; ++r) rp32[r] = (int32_t)c->rowptr[r];
  // upload
  dfree(c->d_nodeInfo); dfree(c->d_colofs); dfree(c->d_sconn); dfree(c->d_elemByColor); dfree(c->d_rowptr32);
  dfree(c->d_colidx); dfree(c->d_assy); dfree(c->d_vals); dfree(c->d_vals32); dfree(c->d_rhs); dfree(c->d_du);
  dfree(c->d_blkofs); dfree(c->d_natofs); dfree(c->d_blkPtr); dfree(c->d_kb); dfree(c->d_kb32);
  if (c->la) la_destroy(c);
  // the Hexa8 loop assembles the node-pair block form
  c->kb_on = c->et == HEXA8;
  for (int m = 0; m < 3; ++m) c->asm_variant[m] = -1;
  c->csr_version = -1;
  int rc;
  if ((rc = dalloc(&c->d_nodeInfo, info.size())) || (rc = dalloc(&c->d_colofs, scolofs.size())) ||
      (rc = dalloc(&c->d_sconn, sconn.size())) ||
      (rc = dalloc(&c->d_elemByColor, byColor.size())) || (rc = dalloc(&c->d_rowptr32, rp32.size())) ||
      (rc = dalloc(&c->d_colidx, c->colidx.size())) || (rc = dalloc(&c->d_assy, c->assyForSoln.size())) ||
      (rc = dalloc(&c->d_vals, c->nnz)) || (rc = dalloc(&c->d_rhs, nf)) || (rc = dalloc(&c->d_du, nf)))
    return rc;
  HIP_OK(hipMemcpy(c->d_nodeInfo, info.data(), info.size() * sizeof(int32_t), hipMemcpyHostToDevice));
  HIP_OK(hipMemcpy(c->d_colofs, scolofs.data(), scolofs.size() * sizeof(int32_t), hipMemcpyHostToDevice));
  if (c->kb_on) {
    if ((rc = dalloc(&c->d_blkofs, sblkofs.size())) || (rc = dalloc(&c->d_natofs, natofs.size())) ||
        (rc = dalloc(&c->d_blkPtr, c->blkPtr.size())) || (rc = dalloc(&c->d_kb, std::max<int64_t>(c->nbpairs, 1) * 16)))
      return rc;
    HIP_OK(hipMemcpy(c->d_blkofs, sblkofs.data(), sblkofs.size() * sizeof(int32_t), hipMemcpyHostToDevice));
    HIP_OK(hipMemcpy(c->d_natofs, natofs.data(), natofs.size() * sizeof(int32_t), hipMemcpyHostToDevice));
    HIP_OK(hipMemcpy(c->d_blkPtr, c->blkPtr.data(), c->blkPtr.size() * sizeof(int64_t), hipMemcpyHostToDevice));
    HIP_OK(hipMemset(c->d_kb, 0, c->nbpairs * 16 * sizeof(double)));
  }
  HIP_OK(hipMemcpy(c->d_sconn, sconn.data(), sconn.size() * sizeof(int32_t), hipMemcpyHostToDevice));
  HIP_OK(hipMemcpy(c->d_elemByColor, byColor.data(), byColor.size() * sizeof(int32_t), hipMemcpyHostToDevice));
  HIP_OK(hipMemcpy(c->d_rowptr32, rp32.data(), rp32.size() * sizeof(int32_t), hipMemcpyHostToDevice));
  HIP_OK(hipMemcpy(c->d_colidx, c->colidx.data(), c->colidx.size() * sizeof(int32_t), hipMemcpyHostToDevice));
  if (nf)
    HIP_OK(hipMemcpy(c->d_assy, c->assyForSoln.data(), nf * sizeof(int32_t), hipMemcpyHostToDevice));
  HIP_OK(hipMemset(c->d_vals, 0, c->nnz * sizeof(double)));
  HIP_OK(hipMemset(c->d_rhs, 0, std::max<int64_t>(nf, 1) * sizeof(double)));
  HIP_OK(hipMemset(c->d_du, 0, std::max<int64_t>(nf, 1) * sizeof(double)));
  if ((rc = la_create(c))) return rc;
  c->status |= ST_PATTERN;
  c->status &= ~ST_ASSEMBLED;
  if (nfree_out) *nfree_out = nf;
  if (nnz_out) *nnz_out = c->nnz;
  return SFEM_OK;
}

int sfem_get_pattern(sfem_ctx* c, int64_t* rowptr, int32_t* colidx, int32_t* assy) {
  if (!c) return set_error(SFEM_ERR_ARG, "null context");
  if (!(c->status & ST_PATTERN)) return set_error(SFEM_ERR_STATE, "pattern not built");
  if (rowptr) std::memcpy(rowptr, c->rowptr.data(), (c->nfree + 1) * sizeof(int64_t));
  if (colidx) std::memcpy(colidx, c->colidx.data(), c->nnz * sizeof(int32_t));
  if (assy) std::memcpy(assy, c->assyForSoln.data(), c->nfree * sizeof(int32_t));
  return SFEM_OK;
}

int sfem_set_params(sfem_ctx* c, const double* elemData, int tis, double rhoInf, double dt) {
  if (!c || !elemData) return set_error(SFEM_ERR_ARG, "null argument");
  double td[16];
  if (time_params(tis, rhoInf, dt, td) != 0) return set_error(SFEM_ERR_ARG, "invalid tis");
  if (!(elemData[0] > 0.0) || !(elemData[1] > 0.0)) return set_error(SFEM_ERR_ARG, "rho and mu must be positive");
  std::memcpy(c->elemData, elemData, sizeof(c->elemData));
  c->tis = tis;
  c->rhoInf = rhoInf;
  c->dt = dt;
  std::memcpy(c->td, td, sizeof(td));
  c->status |= ST_PARAMS;
  return SFEM_OK;
}

int sfem_set_precision(sfem_ctx* c, int precision) {
  if (!c) return set_error(SFEM_ERR_ARG, "null context");
  if (precision != SFEM_PREC_FP64 && precision != SFEM_PREC_MIXED && precision != SFEM_PREC_STORE32)
    return set_error(SFEM_ERR_ARG, "bad precision");
  HIP_OK(hipSetDevice(c->device));
  if (precision == SFEM_PREC_MIXED && !c->d_vals32 && (c->status & ST_PATTERN)) {
    int rc = dalloc(&c->d_vals32, c->nnz);
    if (rc) return rc;
    HIP_OK(hipMemset(c->d_vals32, 0, c->nnz * sizeof(float)));
  }
  c->precision = precision;
  c->status &= ~ST_ASSEMBLED;
  return SFEM_OK;
}

int sfem_set_predictor(sfem_ctx* c, int predictor) {
  if (!c) return set_error(SFEM_ERR_ARG, "null context");
  if (predictor != SFEM_PRED_PREVIOUS && predictor != SFEM_PRED_EXTRAPOLATE && predictor != SFEM_PRED_LEARNED &&
      predictor != SFEM_PRED_QUADRATIC)
    return set_error(SFEM_ERR_ARG, "bad predictor");
  if (predictor == SFEM_PRED_LEARNED && !c->learner) return set_error(SFEM_ERR_STATE, "attach a learner first");
  c->predictor = predictor;
  return SFEM_OK;
}

int sfem_predictor_stats(sfem_ctx* c, int64_t* counts, double* last_norms) {
  if (!c) return set_error(SFEM_ERR_ARG, "null context");
  for (int k = 0; k < 3; ++k) {
    if (counts) counts[k] = c->pred_count[k];
    if (last_norms) last_norms[k] = c->pred_norms[k];
  }
  return SFEM_OK;
}

int sfem_set_inner_tolerance(sfem_ctx* c, double theta, double rtol_max) {
  if (!c) return set_error(SFEM_ERR_ARG, "null context");
  if (!(theta >= 0.0) || !std::isfinite(theta)) return set_error(SFEM_ERR_ARG, "theta must be finite and >= 0");
  if (!(rtol_max > 0.0 && rtol_max < 1.0)) return set_error(SFEM_ERR_ARG, "rtol_max must lie in (0, 1)");
  c->inner_theta = theta;
  c->inner_rtol_max = rtol_max;
  return SFEM_OK;
}

int sfem_set_pc_lag(sfem_ctx* c, int lag) {
  if (!c) return set_error(SFEM_ERR_ARG, "null context");
  if (lag < 1) return set_error(SFEM_ERR_ARG, "pc lag must be >= 1");
  if (!c->la) return set_error(SFEM_ERR_STATE, "build the pattern first");
  c->la->pc_lag = lag;
  c->la->lag_count = 0;
  return SFEM_OK;
}

int sfem_set_coarse_correction(sfem_ctx* c, int k, int mult) {
  if (!c) return set_error(SFEM_ERR_ARG, "null context");
  if (!c->la) return set_error(SFEM_ERR_STATE, "build the pattern first");
  if (k < 0) return set_error(SFEM_ERR_ARG, "boxes per direction must be >= 0");
  int rc = sfem::coarse_set(c, k, mult);
  if (rc) return rc;
  c->la->pc_K_version = -1;  // the next solve builds the preconditioner again
  return SFEM_OK;
}

int sfem_coarse_info(sfem_ctx* c, int64_t* info) {
  if (!c || !info) return set_error(SFEM_ERR_ARG, "null argument");
  if (!c->la) return set_error(SFEM_ERR_STATE, "build the pattern first");
  return sfem::coarse_info(c, info);
}

int sfem_coarse_inverse(sfem_ctx* c, int which, double* out, int cap) {
  if (!c || !out || (which != 0 && which != 1)) return set_error(SFEM_ERR_ARG, "bad argument");
  if (!c->la) return set_error(SFEM_ERR_STATE, "build the pattern first");
  return sfem::coarse_get_inverse(c, which, out, cap);
}

int sfem_coarse_step(sfem_ctx* c, const sfem_solver_opts* o) {
  if (!c || !o) return set_error(SFEM_ERR_ARG, "null argument");
  if (!(c->status & ST_ASSEMBLED)) return set_error(SFEM_ERR_STATE, "assemble the system first");
  HIP_OK(hipSetDevice(c->device));
  return sfem::la_coarse_step(c, o);
}

static int run_element_loop(sfem_ctx* c, int mode, int lift, int variant = -1, bool timed = true);

// ||F_free|| at the state x with the prescribed values of this step in place (x is
// overwritten at the prescribed DOFs).  The residual is affine in the state (the
// tangent and the stabilisation use only previous steps), so this is the
// iteration-0 norm of a Newton loop started from x.
static int residual_norm_at(sfem_ctx* c, double* x, double timeFact, double* nrm) {
  const int64_t n = nstate(c);
  HIP_OK(launch_set_bc(c->stream, (int64_t)c->dbc_ind.size(), c->d_dbc_ind, c->d_dbc_val, timeFact, x));
  HIP_OK(launch_iter_step(c->stream, n, time_coeffs(c), x, c->d_prev, c->d_prev2, c->d_prev3, c->d_prev4,
                          c->d_dot, c->d_dotPrev, c->d_cur, c->d_dotCur));
  HIP_OK(launch_zero_f64(c->stream, c->nfree, c->d_rhs));
  int rc = run_element_loop(c, 1, 0);
  if (rc) return rc;
  rc = la_norm2(c, c->d_rhs, nrm);
  c->kt.resolve();  // la_norm2 drained the stream
  return rc;
}

// d_soln holds the linearly extrapolated start; replace it by the quadratically
// extrapolated or the learned start, whichever has the smallest residual norm.
static int judge_learned_start(sfem_ctx* c, double timeFact) {
  if (!c->learner || !c->d_cand) return set_error(SFEM_ERR_STATE, "no learner attached");
  const int64_t n = nstate(c);
  double* cand[3] = {c->d_soln, c->d_cand + n, c->d_cand};  // linear, quadratic, learned
  int rc;
  HIP_OK(launch_extrapolate2(c->stream, n, c->d_prev, c->d_prev2, c->d_prev3, cand[1]));
  if ((rc = learned_candidate(c, c->d_prev, c->d_prev2, c->d_prev3, cand[2]))) return rc;
  int best = 0;
  for (int k = 0; k < 3; ++k) {
    if ((rc = residual_norm_at(c, cand[k], timeFact, &c->pred_norms[k]))) return rc;
    if (k > 0 && std::isfinite(c->pred_norms[k]) && c->pred_norms[k] < c->pred_norms[best]) best = k;
  }
  if (best > 0)
    HIP_OK(hipMemcpyAsync(c->d_soln, cand[best], n * sizeof(double), hipMemcpyDeviceToDevice, c->stream));
  ++c->pred_count[best];
  return SFEM_OK;
}

int sfem_time_update(sfem_ctx* c, double timeNow) {
  if (!c) return set_error(SFEM_ERR_ARG, "null context");
  // the lid ramp of StabFEM2.cpp:636-644 (a ramp factor <= 1 passes through unchanged)
  const double timeFact = timeNow <= 1.0 ? timeNow : 1.0;
  if (!(c->status & ST_PARAMS)) return set_error(SFEM_ERR_STATE, "parameters not set");
  HIP_OK(hipSetDevice(c->device));
  time_params(c->tis, c->rhoInf, c->dt, c->td);  // setTimeParam each step (StabFEM2.cpp:632)
  const int64_t n = nstate(c);
  HIP_OK(launch_time_update(c->stream, n, c->d_soln, c->d_prev, c->d_prev2, c->d_prev3, c->d_prev4, c->d_dot,
                            c->d_dotPrev));
  if (c->predictor == SFEM_PRED_EXTRAPOLATE || c->predictor == SFEM_PRED_LEARNED)
    HIP_OK(launch_extrapolate(c->stream, n, c->d_prev, c->d_prev2, c->d_soln));
  if (c->predictor == SFEM_PRED_QUADRATIC)
    HIP_OK(launch_extrapolate2(c->stream, n, c->d_prev, c->d_prev2, c->d_prev3, c->d_soln));
  if (c->predictor == SFEM_PRED_LEARNED) {
    int rc = judge_learned_start(c, timeFact);
    if (rc) return rc;
  }
  HIP_OK(launch_assign_bc(c->stream, (int64_t)c->dbc_ind.size(), c->d_dbc_ind, c->d_dbc_val, timeFact, c->d_soln,
                          c->d_applied));
  return SFEM_OK;
}

static int run_element_loop(sfem_ctx* c, int mode, int lift, int variant, bool timed) {
  const AsmView v0 = asm_view(c);
  const ElemParams p = elem_params(c);
  const int npe = c->dims.npe;
  if (variant < 0) variant = std::max(0, c->asm_variant[mode]);
  for (int k = 0; k < c->ncolors; ++k) {
    const int64_t n = c->colorPtr[k + 1] - c->colorPtr[k];
    AsmView v = v0;  // slot-ordered arrays start at the colour's first slot
    v.colofs = v0.colofs + (int64_t)c->colorPtr[k] * npe * npe;
    if (v0.blkofs) v.blkofs = v0.blkofs + (int64_t)c->colorPtr[k] * npe * npe;
    v.sconn = v0.sconn + (int64_t)c->colorPtr[k] * npe;
    if (timed) c->kt.begin(SFEM_KT_ASSEMBLY, c->stream);
    HIP_OK(launch_assemble(c->et, c->stream, v, p, c->d_elemByColor + c->colorPtr[k], n, mode, lift, -1, variant));
    if (timed) c->kt.end(SFEM_KT_ASSEMBLY, c->stream);
  }
  return SFEM_OK;
}

// Element-loop variant of an assembly mode: the fastest of the variants timed over the
// whole colour loop (each run starts from cleared outputs; the caller clears and
// assembles again), or variant 0 under fixed tuning and for Hexa8.
static int choose_element_variant(sfem_ctx* c, int mode, int lift) {
  if (c->asm_variant[mode] >= 0) return SFEM_OK;
  if (c->et == HEXA8 || !c->la || c->la->tune_mode == SFEM_TUNE_FIXED) {
    c->asm_variant[mode] = 0;
    return SFEM_OK;
  }
  hipEvent_t e0 = nullptr, e1 = nullptr;
  HIP_OK(hipEventCreate(&e0));
  HIP_OK(hipEventCreate(&e1));
  int best = 0;
  float best_ms = 0.f;
  for (int var = 0; var < kAsmVariants; ++var) {
    HIP_OK(launch_zero_f64(c->stream, c->nfree, c->d_rhs));
    if (mode == 0) HIP_OK(launch_zero_f64(c->stream, c->nnz, c->d_vals));
    if (mode == 2) HIP_OK(launch_zero_f32(c->stream, c->nnz, c->d_vals32));
    HIP_OK(hipEventRecord(e0, c->stream));
    int rc = run_element_loop(c, mode, lift, var, false);
    if (rc) return rc;
    HIP_OK(hipEventRecord(e1, c->stream));
    HIP_OK(hipEventSynchronize(e1));
    float ms = 0.f;
    HIP_OK(hipEventElapsedTime(&ms, e0, e1));
    if (var == 0 || ms < best_ms) { best = var; best_ms = ms; }
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  c->asm_variant[mode] = best;
  return SFEM_OK;
}

// The residual-only assembly of a Newton iteration after the first, as kernel launches
// only (no events, no host synchronisation), so that it can be captured in a graph:
// iteration state, cleared RHS, the element loop.  asm_prepare_residual chooses the
// element-loop variant first (it times variants and synchronises).
}  // extern "C"

namespace sfem {
int asm_prepare_residual(sfem_ctx* c) { return choose_element_variant(c, 1, 0); }

int asm_residual_kernels(sfem_ctx* c) {
  HIP_OK(launch_iter_step(c->stream, nstate(c), time_coeffs(c), c->d_soln, c->d_prev, c->d_prev2, c->d_prev3,
                          c->d_prev4, c->d_dot, c->d_dotPrev, c->d_cur, c->d_dotCur));
  HIP_OK(launch_zero_f64(c->stream, c->nfree, c->d_rhs));
  return run_element_loop(c, 1, 0, -1, false);
}
}  // namespace sfem

extern "C" {

int sfem_assemble(sfem_ctx* c, int iter, int want_K, double* rhs_norm) {
  if (!c) return set_error(SFEM_ERR_ARG, "null context");
  const int need = ST_PREP | ST_PATTERN | ST_PARAMS;
  if ((c->status & need) != need) return set_error(SFEM_ERR_STATE, "prepare, pattern and params must be set");
  if (iter == 0) want_K = 1;
  HIP_OK(hipSetDevice(c->device));
  const int64_t n = nstate(c);
  const bool mixed = c->precision == SFEM_PREC_MIXED;
  if (mixed && !c->d_vals32) {
    int rc = dalloc(&c->d_vals32, c->nnz);
    if (rc) return rc;
  }
  if (mixed && c->kb_on && !c->d_kb32) {
    int rc = dalloc(&c->d_kb32, std::max<int64_t>(c->nbpairs, 1) * 16);
    if (rc) return rc;
  }
  const int64_t nd = (int64_t)c->dbc_ind.size();
  if (mixed && iter == 0)  // prescribed values first; the FP64 residual then needs no lift
    HIP_OK(launch_add_bc(c->stream, nd, c->d_dbc_ind, c->d_dbc_mult, c->d_applied, c->d_soln));
  HIP_OK(launch_iter_step(c->stream, n, time_coeffs(c), c->d_soln, c->d_prev, c->d_prev2, c->d_prev3, c->d_prev4,
                          c->d_dot, c->d_dotPrev, c->d_cur, c->d_dotCur));
  const int mode0 = want_K ? (mixed ? 2 : 0) : 1;
  if (c->asm_variant[mode0] < 0) {
    int rc0 = choose_element_variant(c, mode0, (!mixed && iter == 0) ? 1 : 0);
    if (rc0) return rc0;
  }
  HIP_OK(hipEventRecord(c->ev0, c->stream));
  HIP_OK(launch_zero_f64(c->stream, c->nfree, c->d_rhs));
  int mode;
  if (want_K) {
    if (c->kb_on) {
      // the block form is written whole by its first elements (no clearing)
    } else if (mixed) {
      HIP_OK(launch_zero_f32(c->stream, c->nnz, c->d_vals32));
    } else {
      HIP_OK(launch_zero_f64(c->stream, c->nnz, c->d_vals));
    }
    mode = mixed ? 2 : 0;
  } else {
    mode = 1;
  }
  int rc = run_element_loop(c, mode, (!mixed && iter == 0) ? 1 : 0);
  if (rc) return rc;
  HIP_OK(hipEventRecord(c->ev1, c->stream));
  if (!mixed && iter == 0)
    HIP_OK(launch_add_bc(c->stream, nd, c->d_dbc_ind, c->d_dbc_mult, c->d_applied, c->d_soln));
  double nrm = 0.0;
  if ((rc = la_norm2(c, c->d_rhs, &nrm))) return rc;
  float ms = 0.f;
  HIP_OK(hipEventElapsedTime(&ms, c->ev0, c->ev1));
  c->last_asm_ms = ms;
  c->kt.resolve();  // la_norm2 drained the stream
  if (want_K) {
    c->la_K_version++;
    c->kb_fp32 = mixed;
  }
  if (rhs_norm) *rhs_norm = nrm;
  c->status |= ST_ASSEMBLED;
  return SFEM_OK;
}

int sfem_element_KF(sfem_ctx* c, int64_t e, double* Ke, double* Fe) {
  if (!c) return set_error(SFEM_ERR_ARG, "null context");
  if (e < 0 || e >= c->nElem) return set_error(SFEM_ERR_ARG, "element index out of range");
  const int need = ST_PREP | ST_PATTERN | ST_PARAMS;
  if ((c->status & need) != need) return set_error(SFEM_ERR_STATE, "prepare, pattern and params must be set");
  HIP_OK(hipSetDevice(c->device));
  const int mode = c->precision == SFEM_PREC_MIXED ? 2 : 0;
  HIP_OK(launch_assemble(c->et, c->stream, asm_view(c), elem_params(c), nullptr, 1, mode, 0, e));
  const int nsz = c->dims.npe * c->dims.ndof;
  if (Ke) HIP_OK(hipMemcpyAsync(Ke, c->d_probeK, nsz * nsz * sizeof(double), hipMemcpyDeviceToHost, c->stream));
  if (Fe) HIP_OK(hipMemcpyAsync(Fe, c->d_probeF, nsz * sizeof(double), hipMemcpyDeviceToHost, c->stream));
  HIP_OK(hipStreamSynchronize(c->stream));
  return SFEM_OK;
}

int sfem_get_system(sfem_ctx* c, double* vals, double* rhs) {
  if (!c) return set_error(SFEM_ERR_ARG, "null context");
  if (!(c->status & ST_ASSEMBLED)) return set_error(SFEM_ERR_STATE, "nothing assembled");
  HIP_OK(hipSetDevice(c->device));
  if (vals) {
    int rc = la_ensure_csr(c);
    if (rc) return rc;
  }
  HIP_OK(hipStreamSynchronize(c->stream));
  if (vals) {
    if (c->precision == SFEM_PREC_MIXED) {
      std::vector<float> tmp(c->nnz);
      HIP_OK(hipMemcpy(tmp.data(), c->d_vals32, c->nnz * sizeof(float), hipMemcpyDeviceToHost));
      for (int64_t k = 0; k < c->nnz; ++k) vals[k] = tmp[k];
    } else {
      HIP_OK(hipMemcpy(vals, c->d_vals, c->nnz * sizeof(double), hipMemcpyDeviceToHost));
    }
  }
  if (rhs) HIP_OK(hipMemcpy(rhs, c->d_rhs, c->nfree * sizeof(double), hipMemcpyDeviceToHost));
  return SFEM_OK;
}

int sfem_solve(sfem_ctx* c, const sfem_solver_opts* o, int* iters, double* relres) {
  if (!c || !o) return set_error(SFEM_ERR_ARG, "null argument");
  if (!(c->status & ST_ASSEMBLED)) return set_error(SFEM_ERR_STATE, "assemble the system first");  // SolverPetsc.cpp:288-292
  if (o->rtol <= 0.0 || o->max_it <= 0 || !(o->atol >= 0.0)) return set_error(SFEM_ERR_ARG, "bad solver options");
  if (o->ksp_type != SFEM_KSP_BCGS && o->ksp_type != SFEM_KSP_GMRES && o->ksp_type != SFEM_KSP_PREONLY)
    return set_error(SFEM_ERR_ARG, "bad ksp_type");
  if (o->pc_type < SFEM_PC_NONE || o->pc_type > SFEM_PC_LU) return set_error(SFEM_ERR_ARG, "bad pc_type");
  if ((o->ksp_type == SFEM_KSP_PREONLY) != (o->pc_type == SFEM_PC_LU))
    return set_error(SFEM_ERR_UNSUPPORTED, "preonly runs with the LU direct solver, and LU only as preonly");
  HIP_OK(hipSetDevice(c->device));
  int it = 0;
  double rr = 0.0;
  HIP_OK(hipEventRecord(c->ev2, c->stream));
  int rc = la_solve(c, o, &it, &rr);
  HIP_OK(hipEventRecord(c->ev3, c->stream));
  HIP_OK(hipEventSynchronize(c->ev3));
  float ms = 0.f;
  HIP_OK(hipEventElapsedTime(&ms, c->ev2, c->ev3));
  c->last_solve_ms = ms;
  c->kt.resolve();
  if (iters) *iters = it;
  if (relres) *relres = rr;
  return rc;
}

int sfem_direct_info(sfem_ctx* c, int* kl, int* ku) {
  if (!c || !kl || !ku) return set_error(SFEM_ERR_ARG, "null argument");
  if (!c->la) { *kl = *ku = 0; return SFEM_OK; }
  return la_direct_info(c, kl, ku);
}

int sfem_setup_solver(sfem_ctx* c, const sfem_solver_opts* o) {
  if (!c || !o) return set_error(SFEM_ERR_ARG, "null argument");
  if (!(c->status & ST_PATTERN)) return set_error(SFEM_ERR_STATE, "pattern not built");
  HIP_OK(hipSetDevice(c->device));
  return la_prepare(c, o);
}

int sfem_update_solution(sfem_ctx* c) {
  if (!c) return set_error(SFEM_ERR_ARG, "null context");
  if (!(c->status & ST_PATTERN)) return set_error(SFEM_ERR_STATE, "pattern not built");
  HIP_OK(hipSetDevice(c->device));
  HIP_OK(launch_update_solution(c->stream, c->nfree, c->d_assy, c->d_du, c->d_soln));
  return SFEM_OK;
}

int sfem_step(sfem_ctx* c, double timeNow, const sfem_solver_opts* o, double newton_tol, int max_iter,
              sfem_step_stats* stats) {
  if (!c || !o) return set_error(SFEM_ERR_ARG, "null argument");
  sfem_step_stats st;
  std::memset(&st, 0, sizeof st);
  int rc = sfem_time_update(c, timeNow);
  if (rc) return rc;
  bool conv = false;
  for (int iter = 0; iter < max_iter; ++iter) {
    double nrm = 0.0;
    // the tangent depends only on the previous step's state, so it is assembled once per step
    if ((rc = sfem_assemble(c, iter, iter == 0 ? 1 : 0, &nrm))) return rc;
    st.t_assembly_ms += c->last_asm_ms;
    if (iter == 0) st.rhs_norm0 = nrm;
    st.rhs_norm = nrm;
    st.newton_iters = iter + 1;
    if (!std::isfinite(nrm)) { if (stats) *stats = st; return set_error(SFEM_ERR_DIVERGED, "non-finite residual"); }
    if (nrm < newton_tol) { conv = true; break; }
    if (c->precision == SFEM_PREC_MIXED && c->la->coarse_k > 0 && o->pc_type == SFEM_PC_ILU0) {
      // the coarse step in FP64 (x += Z E^-1 Z^T F), then the residual again: the FP32
      // inner solve below leaves the global constant pressure mode to it
      if ((rc = sfem_coarse_step(c, o)) || (rc = sfem_update_solution(c))) { if (stats) *stats = st; return rc; }
      if ((rc = sfem_assemble(c, iter + 1, 0, &nrm))) return rc;
      st.t_assembly_ms += c->last_asm_ms;
      st.rhs_norm = nrm;
      if (!std::isfinite(nrm)) { if (stats) *stats = st; return set_error(SFEM_ERR_DIVERGED, "non-finite residual"); }
      if (nrm < newton_tol) { conv = true; break; }
    }
    int its = 0;
    double rr = 0.0;
    sfem_solver_opts oi = *o;
    if (c->inner_theta > 0.0 && c->precision != SFEM_PREC_FP64 && nrm > 0.0)
      oi.rtol = std::min(c->inner_rtol_max, std::max(o->rtol, c->inner_theta * newton_tol / nrm));
    rc = sfem_solve(c, &oi, &its, &rr);
    st.krylov_iters += its;
    st.t_solver_ms += c->last_solve_ms;
    if (rc) { if (stats) *stats = st; return rc; }
    if ((rc = sfem_update_solution(c))) return rc;
  }
  if (stats) *stats = st;
  if (!conv) return set_error(SFEM_ERR_NOTCONV, "Newton loop reached its iteration cap");
  return SFEM_OK;
}

int sfem_get_solution(sfem_ctx* c, double* soln) {
  if (!c || !soln) return set_error(SFEM_ERR_ARG, "null argument");
  HIP_OK(hipSetDevice(c->device));
  HIP_OK(hipStreamSynchronize(c->stream));
  HIP_OK(hipMemcpy(soln, c->d_soln, nstate(c) * sizeof(double), hipMemcpyDeviceToHost));
  return SFEM_OK;
}

int sfem_set_solution(sfem_ctx* c, const double* soln) {
  if (!c || !soln) return set_error(SFEM_ERR_ARG, "null argument");
  HIP_OK(hipSetDevice(c->device));
  HIP_OK(hipStreamSynchronize(c->stream));
  HIP_OK(hipMemcpy(c->d_soln, soln, nstate(c) * sizeof(double), hipMemcpyHostToDevice));
  return SFEM_OK;
}

int sfem_get_increment(sfem_ctx* c, double* du) {
  if (!c || !du) return set_error(SFEM_ERR_ARG, "null argument");
  if (!(c->status & ST_PATTERN)) return set_error(SFEM_ERR_STATE, "pattern not built");
  HIP_OK(hipSetDevice(c->device));
  HIP_OK(hipStreamSynchronize(c->stream));
  HIP_OK(hipMemcpy(du, c->d_du, c->nfree * sizeof(double), hipMemcpyDeviceToHost));
  return SFEM_OK;
}

int sfem_set_state(sfem_ctx* c, const double* soln, const double* prev, const double* cur,
                   const double* dotCur, const double* dot, const double* dotPrev) {
  if (!c) return set_error(SFEM_ERR_ARG, "null context");
  HIP_OK(hipSetDevice(c->device));
  HIP_OK(hipStreamSynchronize(c->stream));
  const size_t b = nstate(c) * sizeof(double);
  if (soln) HIP_OK(hipMemcpy(c->d_soln, soln, b, hipMemcpyHostToDevice));
  if (prev) HIP_OK(hipMemcpy(c->d_prev, prev, b, hipMemcpyHostToDevice));
  if (cur) HIP_OK(hipMemcpy(c->d_cur, cur, b, hipMemcpyHostToDevice));
  if (dotCur) HIP_OK(hipMemcpy(c->d_dotCur, dotCur, b, hipMemcpyHostToDevice));
  if (dot) HIP_OK(hipMemcpy(c->d_dot, dot, b, hipMemcpyHostToDevice));
  if (dotPrev) HIP_OK(hipMemcpy(c->d_dotPrev, dotPrev, b, hipMemcpyHostToDevice));
  return SFEM_OK;
}

int sfem_reset_state(sfem_ctx* c) {
  if (!c) return set_error(SFEM_ERR_ARG, "null context");
  HIP_OK(hipSetDevice(c->device));
  const int64_t n = nstate(c);
  double* sv[] = {c->d_soln, c->d_prev, c->d_prev2, c->d_prev3, c->d_prev4, c->d_dot,
                  c->d_dotPrev, c->d_cur, c->d_dotCur, c->d_applied};
  for (double* q : sv) HIP_OK(launch_zero_f64(c->stream, n, q));
  HIP_OK(hipStreamSynchronize(c->stream));
  return SFEM_OK;
}

int sfem_enable_kernel_timing(sfem_ctx* c, int enable) {
  if (!c) return set_error(SFEM_ERR_ARG, "null context");
  HIP_OK(hipSetDevice(c->device));
  HIP_OK(hipStreamSynchronize(c->stream));
  c->kt.reset();
  c->kt.on = enable != 0;
  return SFEM_OK;
}

int sfem_kernel_timing(sfem_ctx* c, int kind, int64_t* launches, double* total_ms) {
  if (!c) return set_error(SFEM_ERR_ARG, "null context");
  if (kind < 0 || kind >= SFEM_KT_NKINDS) return set_error(SFEM_ERR_ARG, "bad kernel kind");
  HIP_OK(hipSetDevice(c->device));
  HIP_OK(hipStreamSynchronize(c->stream));
  c->kt.resolve();
  if (launches) *launches = c->kt.count[kind];
  if (total_ms) *total_ms = c->kt.ms[kind];
  return SFEM_OK;
}

int sfem_element_tuning(sfem_ctx* c, int* out) {
  if (!c || !out) return set_error(SFEM_ERR_ARG, "null argument");
  for (int m = 0; m < 3; ++m) out[m] = c->asm_variant[m];
  return SFEM_OK;
}

int sfem_tuning(sfem_ctx* c, int* out) {
  if (!c || !out) return set_error(SFEM_ERR_ARG, "null argument");
  return la_tuning(c, out);
}

int sfem_get_ilu(sfem_ctx* c, int64_t* nnz, int32_t* perm, int32_t* rowptr, int32_t* colidx, double* vals) {
  if (!c) return set_error(SFEM_ERR_ARG, "null context");
  return la_get_ilu(c, nnz, perm, rowptr, colidx, vals);
}

int sfem_tile_info(sfem_ctx* c, int64_t* ntiles, int64_t* ncolors, int64_t* tile_vals) {
  if (!c) return set_error(SFEM_ERR_ARG, "null context");
  return la_tile_info(c, ntiles, ncolors, tile_vals);
}

int sfem_pc_apply(sfem_ctx* c, const double* r, double* z) {
  if (!c || !r || !z) return set_error(SFEM_ERR_ARG, "null argument");
  HIP_OK(hipSetDevice(c->device));
  return la_pc_apply_host(c, r, z);
}

int sfem_pc_apply_noalloc(sfem_ctx* c, const double* r, double* z) {
  if (!c || !r || !z) return set_error(SFEM_ERR_ARG, "null argument");
  HIP_OK(hipSetDevice(c->device));
  return la_pc_apply_noalloc(c, r, z);
}

int sfem_block_info(sfem_ctx* c, int64_t* nblk, int64_t* npairs, int64_t* nvals) {
  if (!c) return set_error(SFEM_ERR_ARG, "null context");
  return la_block_info(c, nblk, npairs, nvals);
}

int sfem_solver_stats(sfem_ctx* c, int64_t* fp32_fallbacks, int64_t* fp32_partial) {
  if (!c || !fp32_fallbacks) return set_error(SFEM_ERR_ARG, "null argument");
  *fp32_fallbacks = c->la ? c->la->fallbacks : 0;
  if (fp32_partial) *fp32_partial = (c->la ? c->la->partial : 0) + batch_stalled(c);
  return SFEM_OK;
}

int sfem_tuning_levels(sfem_ctx* c, int fp32, int* out, int cap, int* nlevels) {
  if (!c || !nlevels || (cap > 0 && !out) || cap < 0) return set_error(SFEM_ERR_ARG, "bad argument");
  if (!c->la) { *nlevels = 0; return SFEM_OK; }
  return la_tuning_levels(c, fp32, out, cap, nlevels);
}

int sfem_set_tuning_mode(sfem_ctx* c, int mode) {
  if (!c) return set_error(SFEM_ERR_ARG, "null context");
  if (!c->la) return set_error(SFEM_ERR_STATE, "build the pattern first");
  return la_set_tuning_mode(c, mode);
}

int sfem_get_tuning_table(sfem_ctx* c, int fp32, int* table, int cap, int* n) {
  if (!c || !n || cap < 0 || (cap > 0 && !table)) return set_error(SFEM_ERR_ARG, "bad argument");
  if (!c->la) return set_error(SFEM_ERR_STATE, "build the pattern first");
  return la_get_tuning_table(c, fp32, table, cap, n);
}

int sfem_set_tuning_table(sfem_ctx* c, int fp32, const int* table, int n) {
  if (!c || !table || n < 0) return set_error(SFEM_ERR_ARG, "bad argument");
  if (!c->la) return set_error(SFEM_ERR_STATE, "build the pattern first");
  return la_set_tuning_table(c, fp32, table, n);
}

int sfem_info(sfem_ctx* c, int64_t* nfree, int64_t* nnz, int* ncolors, int* ndof) {
  if (!c) return set_error(SFEM_ERR_ARG, "null context");
  if (nfree) *nfree = c->nfree;
  if (nnz) *nnz = c->nnz;
  if (ncolors) *ncolors = c->ncolors;
  if (ndof) *ndof = c->dims.ndof;
  return SFEM_OK;
}

int sfem_last_timings(sfem_ctx* c, double* assembly_ms, double* solve_ms) {
  if (!c) return set_error(SFEM_ERR_ARG, "null context");
  if (assembly_ms) *assembly_ms = c->last_asm_ms;
  if (solve_ms) *solve_ms = c->last_solve_ms;
  return SFEM_OK;
}

}  // extern "C"
