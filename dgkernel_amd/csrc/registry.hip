// Kernel registry: the replacement for OCCA's JIT build of user OKL strings
// (runner/runner.go:88-123 BuildKernel -> Device.BuildKernelFromString) and
// for OCCAKernel.RunWithArgs (runner/kernel_execution.go:57).
//
// There is no OKL compiler here. A caller asks for a kernel by the name its
// OKL source declares and the signature string the runner generated for it
// (runner/kernel_config.go:186-246, argument order of
// runner/kernel_execution.go:150-227: K | device matrices sorted by name |
// X_global, X_offsets per array | scalars). The registry parses that
// signature, checks that it carries the arrays the native kernel needs, and
// at launch maps the argument list onto the native kernel by name.
//
// Registered kernels (the hot-path bodies of the reference's callers):
//   "differentiate"  with DuDx/DuDy/DuDz, U, Rx..Tz and Dr_*, Ds_*, Dt_*:
//       tetnudg/tet_nudg_mesh_physical_derivative_test.go:850-892 (FP64
//       metrics) and :240-278 (FP32 metrics/outputs);
//   "differentiate"  with Dx, U, Rx and Dr_*: Dx = Rx .* (Dr U)
//       (same file :92-113);
//   "differentiate"  with Dx, U and Dr_* (no Rx): Dx = Dr U;
//   "matmul"         with one matrix, one input and one output array:
//       OUT = M IN per element (runner/runner_matrix_test.go:13-125).
#include <cstring>
#include <map>
#include <string>

#include "dgk_internal.h"

namespace dgk {
int phys_deriv_nodal_any(dgk_device dev, int N, int npart, const int64_t* K, const double* Dr,
                         const double* Ds, const double* Dt, const double* u, const void* u_off,
                         const void* const* metrics, int mtype, const void* m_off, void* dx,
                         void* dy, void* dz, int otype, const void* o_off, int off_size);
int scaled_matmul_any(dgk_device dev, int rows, int cols, const double* A_colmajor, int npart,
                      const int64_t* K, const double* in, const void* in_off, const void* sc,
                      int sc_type, const void* sc_off, double* out, const void* out_off,
                      int off_size);
int transpose_colmajor(dgk_device dev, int rows, int cols, const double* in, double* out);
}  // namespace dgk

namespace {

enum Kind { K_DERIV3 = 1, K_DX_SCALED = 2, K_DX = 3, K_MATMUL = 4 };

struct SigArg {
  std::string type;  // e.g. "double", "float", "int_t"
  std::string name;
  bool is_const = false, is_ptr = false;
};

std::string trim(const std::string& s) {
  size_t a = s.find_first_not_of(" \t\r\n"), b = s.find_last_not_of(" \t\r\n");
  return a == std::string::npos ? "" : s.substr(a, b - a + 1);
}

bool parse_signature(const char* sig, std::vector<SigArg>& out) {
  std::string s(sig ? sig : "");
  size_t pos = 0;
  while (pos <= s.size()) {
    size_t c = s.find(',', pos);
    std::string tok = trim(s.substr(pos, c == std::string::npos ? std::string::npos : c - pos));
    pos = (c == std::string::npos) ? s.size() + 1 : c + 1;
    if (tok.empty()) continue;
    SigArg a;
    if (tok.rfind("const ", 0) == 0) {
      a.is_const = true;
      tok = trim(tok.substr(6));
    }
    size_t star = tok.find('*');
    a.is_ptr = star != std::string::npos;
    size_t name_start = tok.find_last_of(" *");
    if (name_start == std::string::npos) return false;
    a.name = trim(tok.substr(name_start + 1));
    a.type = trim(tok.substr(0, a.is_ptr ? star : name_start));
    if (a.name.empty() || a.type.empty()) return false;
    out.push_back(a);
  }
  return !out.empty();
}

bool ends_with(const std::string& s, const std::string& suf) {
  return s.size() >= suf.size() && s.compare(s.size() - suf.size(), suf.size(), suf) == 0;
}

int order_of_np(int np) {
  for (int N = 0; N <= 10; ++N)
    if ((N + 1) * (N + 2) * (N + 3) / 6 == np) return N;
  return -1;
}

}  // namespace

struct dgk_kernel_s {
  dgk_device dev = nullptr;
  std::string name;
  Kind kind = K_DERIV3;
  dgk_kernel_consts c{};
  std::vector<SigArg> args;
  std::map<std::string, int> index;             // signature name -> position
  std::map<std::string, std::string> arr_type;  // array base name -> element type
  std::vector<std::string> dev_mats;            // device matrix names in signature order
  struct Static {
    int rows = 0, cols = 0;
    double* rowmajor = nullptr;  // device
    double* colmajor = nullptr;  // device
  };
  std::map<std::string, Static> statics;
  double* scratch = nullptr;  // 3*Np*Np row-major matrices transposed from device args
  std::string mat_prefix[3];  // resolved "Dr_...", "Ds_...", "Dt_..." (or the matmul matrix)
};

namespace {

bool has_array(const dgk_kernel_s* k, const std::string& base) {
  return k->arr_type.count(base) != 0;
}

// Matrix name starting with prefix ("Dr_") among device-matrix args; empty if
// none (then it must come as a static matrix via dgk_kernel_set_matrix).
std::string find_matrix(const dgk_kernel_s* k, const std::string& prefix) {
  for (const auto& m : k->dev_mats)
    if (m.rfind(prefix, 0) == 0) return m;
  return "";
}

}  // namespace

extern "C" int dgk_kernel_list(char* buf, size_t size) {
  const char* names = "differentiate\nmatmul\n";
  DGK_REQUIRE(buf && size > strlen(names), "dgk_kernel_list: buffer too small");
  memcpy(buf, names, strlen(names) + 1);
  return DGK_OK;
}

extern "C" int dgk_kernel_get(dgk_device dev, const char* name, const char* signature,
                              const dgk_kernel_consts* consts, dgk_kernel* out) {
  dgk::clear_error();
  DGK_REQUIRE(dev && name && signature && consts && out, "dgk_kernel_get: NULL argument");
  *out = nullptr;
  DGK_REQUIRE(consts->int_size == 8 || consts->int_size == 4, "dgk_kernel_get: int_size %d",
              consts->int_size);
  DGK_REQUIRE(consts->npart >= 1, "dgk_kernel_get: npart %d", consts->npart);
  std::vector<SigArg> args;
  if (!parse_signature(signature, args)) {
    dgk::set_error("dgk_kernel_get: cannot parse signature \"%s\"", signature);
    return DGK_ERR_INVALID;
  }
  DGK_REQUIRE(args[0].name == "K" && args[0].is_ptr,
              "dgk_kernel_get: signature must start with the K array (runner/kernel_execution.go:153)");
  dgk_kernel k = new dgk_kernel_s();
  k->dev = dev;
  k->name = name;
  k->c = *consts;
  k->args = args;
  for (size_t i = 0; i < args.size(); ++i) {
    k->index[args[i].name] = (int)i;
    const std::string& n = args[i].name;
    if (i == 0) continue;
    if (args[i].is_ptr && ends_with(n, "_global")) {
      k->arr_type[n.substr(0, n.size() - 7)] = args[i].type;
    } else if (args[i].is_ptr && !ends_with(n, "_offsets")) {
      k->dev_mats.push_back(n);
    }
  }
  std::string nm(name);
  int rc = DGK_OK;
  if (nm == "differentiate") {
    bool full = has_array(k, "DuDx") && has_array(k, "DuDy") && has_array(k, "DuDz");
    const char* mets[9] = {"Rx", "Ry", "Rz", "Sx", "Sy", "Sz", "Tx", "Ty", "Tz"};
    for (auto* m : mets) full = full && has_array(k, m);
    full = full && has_array(k, "U");
    if (full) {
      k->kind = K_DERIV3;
    } else if (has_array(k, "Dx") && has_array(k, "U")) {
      k->kind = has_array(k, "Rx") ? K_DX_SCALED : K_DX;
    } else {
      rc = DGK_ERR_NOT_FOUND;
      dgk::set_error("dgk_kernel_get: \"differentiate\" signature lacks U and DuDx/DuDy/DuDz or Dx");
    }
    if (rc == DGK_OK) {
      DGK_REQUIRE(k->arr_type["U"] == "double", "differentiate: U must be double");
      const char* pre[3] = {"Dr_", "Ds_", "Dt_"};
      for (int q = 0; q < 3; ++q) k->mat_prefix[q] = pre[q];
      int np = consts->np;
      if (np <= 0 || order_of_np(np) < 0) {
        rc = DGK_ERR_INVALID;
        dgk::set_error("differentiate: consts.np=%d is not a tetrahedral node count", np);
      } else {
        rc = dgk_malloc(dev, (size_t)3 * np * np * 8, nullptr, (void**)&k->scratch);
      }
    }
  } else if (nm == "matmul") {
    std::string in, outn;
    for (auto& a : k->arr_type) {
      int gi = k->index[a.first + "_global"];
      if (k->args[gi].is_const && in.empty()) in = a.first;
      if (!k->args[gi].is_const && outn.empty()) outn = a.first;
    }
    if (in.empty() || outn.empty() || k->arr_type.size() != 2 || k->dev_mats.size() > 1) {
      rc = DGK_ERR_NOT_FOUND;
      dgk::set_error("dgk_kernel_get: \"matmul\" needs one input array, one output array and one matrix");
    }
    k->kind = K_MATMUL;
    k->mat_prefix[0] = in;    // reuse: input name
    k->mat_prefix[1] = outn;  // output name
    k->mat_prefix[2] = k->dev_mats.empty() ? "" : k->dev_mats[0];
  } else {
    rc = DGK_ERR_NOT_FOUND;
    dgk::set_error("dgk_kernel_get: no native kernel \"%s\" (registry: differentiate, matmul)", name);
  }
  if (rc != DGK_OK) {
    dgk_kernel_free(k);
    return rc;
  }
  *out = k;
  return DGK_OK;
}

extern "C" int dgk_kernel_set_matrix(dgk_kernel k, const char* name, int rows, int cols,
                                     const double* values) {
  dgk::clear_error();
  DGK_REQUIRE(k && name && values && rows > 0 && cols > 0, "dgk_kernel_set_matrix: bad argument");
  std::vector<double> cm((size_t)rows * cols);
  for (int i = 0; i < rows; ++i)
    for (int j = 0; j < cols; ++j) cm[(size_t)j * rows + i] = values[(size_t)i * cols + j];
  dgk_kernel_s::Static st;
  st.rows = rows;
  st.cols = cols;
  int rc = dgk_malloc(k->dev, (size_t)rows * cols * 8, values, (void**)&st.rowmajor);
  if (rc == DGK_OK) rc = dgk_malloc(k->dev, (size_t)rows * cols * 8, cm.data(), (void**)&st.colmajor);
  if (rc != DGK_OK) return rc;
  auto it = k->statics.find(name);
  if (it != k->statics.end()) {
    dgk_free(k->dev, it->second.rowmajor);
    dgk_free(k->dev, it->second.colmajor);
  }
  k->statics[name] = st;
  return DGK_OK;
}

namespace {

const void* mem_arg(const dgk_kernel_s* k, const dgk_arg* args, const std::string& n) {
  auto it = k->index.find(n);
  if (it == k->index.end()) return nullptr;
  if (args[it->second].kind != DGK_ARG_MEM) return nullptr;
  return args[it->second].v.mem;
}

int elem_type(const dgk_kernel_s* k, const std::string& base) {
  auto it = k->arr_type.find(base);
  return (it != k->arr_type.end() && it->second == "float") ? DGK_F32 : DGK_F64;
}

// Row-major device copy of the Np x Np matrix whose name starts with prefix:
// either a device-matrix argument (column-major, transposed into scratch
// slot q) or a static matrix handed over with dgk_kernel_set_matrix.
int matrix_rowmajor(dgk_kernel_s* k, const dgk_arg* args, int q, int np, const double** out) {
  std::string nm = find_matrix(k, k->mat_prefix[q]);
  if (!nm.empty()) {
    const double* colm = (const double*)mem_arg(k, args, nm);
    DGK_REQUIRE(colm, "differentiate: matrix %s is not a memory argument", nm.c_str());
    double* dst = k->scratch + (size_t)q * np * np;
    int rc = dgk::transpose_colmajor(k->dev, np, np, colm, dst);
    if (rc) return rc;
    *out = dst;
    return DGK_OK;
  }
  for (auto& s : k->statics)
    if (s.first.rfind(k->mat_prefix[q], 0) == 0) {
      DGK_REQUIRE(s.second.rows == np && s.second.cols == np, "differentiate: %s is %dx%d, want %dx%d",
                  s.first.c_str(), s.second.rows, s.second.cols, np, np);
      *out = s.second.rowmajor;
      return DGK_OK;
    }
  dgk::set_error("differentiate: no matrix %s* among arguments or static matrices",
                 k->mat_prefix[q].c_str());
  return DGK_ERR_INVALID;
}

}  // namespace

extern "C" int dgk_kernel_launch(dgk_kernel k, const void* host_K, int nargs, const dgk_arg* args) {
  dgk::clear_error();
  DGK_REQUIRE(k && host_K && args, "dgk_kernel_launch: NULL argument");
  DGK_REQUIRE(nargs == (int)k->args.size(), "dgk_kernel_launch: %d args, signature has %d", nargs,
              (int)k->args.size());
  std::vector<int64_t> K(k->c.npart);
  for (int p = 0; p < k->c.npart; ++p)
    K[p] = k->c.int_size == 8 ? ((const int64_t*)host_K)[p] : ((const int32_t*)host_K)[p];
  const int off = k->c.int_size;
  auto G = [&](const std::string& b) { return mem_arg(k, args, b + "_global"); };
  auto O = [&](const std::string& b) { return mem_arg(k, args, b + "_offsets"); };
  const int np = k->c.np;
  switch (k->kind) {
    case K_DERIV3: {
      const double* D[3];
      for (int q = 0; q < 3; ++q) {
        int rc = matrix_rowmajor(k, args, q, np, &D[q]);
        if (rc) return rc;
      }
      const char* mets[9] = {"Rx", "Ry", "Rz", "Sx", "Sy", "Sz", "Tx", "Ty", "Tz"};
      const void* mp[9];
      int mtype = elem_type(k, "Rx");
      for (int q = 0; q < 9; ++q) {
        mp[q] = G(mets[q]);
        DGK_REQUIRE(elem_type(k, mets[q]) == mtype, "differentiate: metric types differ");
      }
      int otype = elem_type(k, "DuDx");
      return dgk::phys_deriv_nodal_any(k->dev, order_of_np(np), k->c.npart, K.data(), D[0], D[1],
                                       D[2], (const double*)G("U"), O("U"), mp, mtype, O("Rx"),
                                       (void*)G("DuDx"), (void*)G("DuDy"), (void*)G("DuDz"), otype,
                                       O("DuDx"), off);
    }
    case K_DX_SCALED:
    case K_DX: {
      DGK_REQUIRE(elem_type(k, "Dx") == DGK_F64, "differentiate: Dx must be double");
      std::string nm = find_matrix(k, "Dr_");
      const double* colm = nullptr;
      if (!nm.empty()) {
        colm = (const double*)mem_arg(k, args, nm);
      } else {
        for (auto& s : k->statics)
          if (s.first.rfind("Dr_", 0) == 0) colm = s.second.colmajor;
      }
      DGK_REQUIRE(colm, "differentiate: no Dr_* matrix");
      bool sc = k->kind == K_DX_SCALED;
      return dgk::scaled_matmul_any(k->dev, np, np, colm, k->c.npart, K.data(), (const double*)G("U"),
                                    O("U"), sc ? G("Rx") : nullptr, elem_type(k, "Rx"),
                                    sc ? O("Rx") : nullptr, (double*)G("Dx"), O("Dx"), off);
    }
    case K_MATMUL: {
      const std::string in = k->mat_prefix[0], outn = k->mat_prefix[1], mn = k->mat_prefix[2];
      const double* colm = nullptr;
      int rows = 0, cols = 0;
      if (!mn.empty()) {
        colm = (const double*)mem_arg(k, args, mn);
        rows = cols = np;
      } else if (k->statics.size() == 1) {
        auto& s = k->statics.begin()->second;
        colm = s.colmajor;
        rows = s.rows;
        cols = s.cols;
      }
      DGK_REQUIRE(colm && rows > 0, "matmul: matrix not available (pass it with dgk_kernel_set_matrix)");
      DGK_REQUIRE(elem_type(k, in) == DGK_F64 && elem_type(k, outn) == DGK_F64,
                  "matmul: arrays must be double");
      return dgk::scaled_matmul_any(k->dev, rows, cols, colm, k->c.npart, K.data(),
                                    (const double*)G(in), O(in), nullptr, DGK_F64, nullptr,
                                    (double*)G(outn), O(outn), off);
    }
  }
  dgk::set_error("dgk_kernel_launch: bad kernel kind");
  return DGK_ERR_INVALID;
}

extern "C" int dgk_kernel_free(dgk_kernel k) {
  if (!k) return DGK_OK;
  for (auto& s : k->statics) {
    dgk_free(k->dev, s.second.rowmajor);
    dgk_free(k->dev, s.second.colmajor);
  }
  dgk_free(k->dev, k->scratch);
  delete k;
  return DGK_OK;
}
