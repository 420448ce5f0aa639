// Source-built kernels: the replacement for OCCA's JIT build of the runner's
// OKL text (runner/runner.go:88-123: BuildKernel -> Device.BuildKernelFromString
// (preamble + "\n" + kernelSource, kernelName, props)) and for
// OCCAKernel.RunWithArgs (runner/kernel_execution.go:57).
//
// There is no OKL compiler on this path. The text is read by a small front end
// that understands what the runner generates and what its callers write:
//
//   preamble (runner/builder/builder.go:172-410)
//     typedef long|int int_t;  #define NPART n;  #define KpartMax n;
//     const double|float NAME[cols][rows] = {...};   static matrices, column-major
//     #define X_PART(part) ((T*)(X_global + X_offsets[part]))
//     #define MATMUL_X(IN, OUT, K_VAL) {...}  and  MATMUL_ADD_X
//   kernel
//     @kernel void name(<signature of runner/kernel_config.go:186-246>) {
//       for (int part = 0; part < NPART; ++part; @outer) { statements } }
//
// with these statements inside the partition loop:
//   [const] T* A = X_PART(part);                      alias of a pooled array
//   MATMUL_X(A, B, K[part]);  MATMUL_ADD_X(A, B, K[part]);
//   element-node loop:   for (k < KpartMax; @inner) if (k < K[part])
//                          for (n < C) { int i = n + k*C; A[i] = expr; ... }
//   flat node loop:      for (i < C*KpartMax; @inner) { A[i] = expr; ... }
//   element matvec loop: for (e < K[part]) for (i < C; @inner) { T s = 0;
//                          for (j < C2) s += M[i*C2 + j] * U[e*C2 + j];
//                          V[e*C + i] = s; }
// Object-like and function-like user macros (#define NP 10) are expanded.
// expr is built from array reads at the node index, scalar arguments and
// numeric literals with + - * / and unary minus, evaluated with C's types
// (float op float rounds to FP32). Anything else is rejected with
// DGK_ERR_UNSUPPORTED naming the source line.
//
// Each statement runs as one launch over all partitions, in order, on
// hand-written kernels: MATMUL on the typed per-element kernel of deriv.hip,
// node loops on a bytecode interpreter kernel (below). The reference's
// "differentiate" body -- MATMUL of three operators into three temporaries,
// then the chain rule with nine metric arrays
// (tetnudg/tet_nudg_mesh_physical_derivative_test.go:850-892, :240-278) -- is
// recognised and runs as one launch of the nodal physical-derivative kernel,
// which still writes the temporaries.
//
// Flat node loops are bounded by each partition's own K[part]*C values (the
// OKL loop bound C*KpartMax reaches past the end of smaller partitions).
#include <cctype>
#include <cmath>
#include <cstdarg>
#include <cstdlib>
#include <cstring>
#include <functional>
#include <map>
#include <memory>
#include <set>
#include <string>
#include <vector>

#include "parts.h"

namespace dgk {
int phys_deriv_nodal_full(dgk_device dev, int N, int npart, const int64_t* K, const void* kdev,
                          int ksize, const double* const* D, int drs, int dcs, const double* u,
                          const void* u_off, const void* const* metrics, const void* const* m_off,
                          int mtype, void* const* outs, const void* const* o_off, int otype,
                          double* const* temps, const void* const* t_off, int off_size,
                          bool force2d);
int matmul_typed(dgk_device dev, int npart, const int64_t* K, const void* kdev, int ksize, int rows,
                 int cols, const void* A, int rs, int cs, int mtype, const void* in,
                 const void* in_off, int itype, void* out, const void* out_off, int otype,
                 int atype, int add, int off_size);
}  // namespace dgk

// ---------------------------------------------------------------------------
// Pointwise interpreter kernel
// ---------------------------------------------------------------------------
namespace {

constexpr int kPwMaxOps = 96;
constexpr int kPwMaxImm = 32;
constexpr int kPwMaxArr = 16;
constexpr int kPwMaxDepth = 12;

enum PwCode : uint8_t { PW_LOAD, PW_IMM, PW_ADD, PW_SUB, PW_MUL, PW_DIV, PW_NEG, PW_STORE };
enum PwFlag : uint8_t { PWF_F32 = 1, PWF_INT = 2 };
enum ElemType { ET_F64 = 0, ET_F32 = 1, ET_I32 = 2, ET_I64 = 3 };

struct PwOp {
  uint8_t code;
  uint8_t flags;
  uint16_t a;
};

struct PwProg {
  int nops, per_elem, off_size, pad;
  PwOp op[kPwMaxOps];
  double imm[kPwMaxImm];
  const void* base[kPwMaxArr];
  const void* off[kPwMaxArr];
  int type[kPwMaxArr];
};

__device__ __forceinline__ double pw_load(int t, const void* b, int64_t i) {
  switch (t) {
    case ET_F64: return ((const double*)b)[i];
    case ET_F32: return (double)((const float*)b)[i];
    case ET_I32: return (double)((const int32_t*)b)[i];
    default: return (double)((const int64_t*)b)[i];
  }
}

__device__ __forceinline__ void pw_store(int t, void* b, int64_t i, double v) {
  switch (t) {
    case ET_F64: ((double*)b)[i] = v; break;
    case ET_F32: ((float*)b)[i] = (float)v; break;
    case ET_I32: ((int32_t*)b)[i] = (int32_t)v; break;
    default: ((int64_t*)b)[i] = (int64_t)v; break;
  }
}

__device__ __forceinline__ double pw_round(uint8_t f, double r) {
  if (f & PWF_F32) r = (double)(float)r;  // float op float: the FP32 result
  if (f & PWF_INT) r = trunc(r);           // int / int truncates
  return r;
}

// One thread per node value; the program is uniform across the block, the
// value stack lives in LDS (stk[depth][thread]).
__global__ void __launch_bounds__(256) pointwise_kernel(dgk::PartTiles pt, int per_tile, PwProg pg) {
  __shared__ double stk[kPwMaxDepth][256];
  int part, first, count;
  if (!dgk::find_tile(pt, per_tile, part, first, count)) return;
  const int tid = threadIdx.x;
  const int64_t n = (int64_t)count * pg.per_elem, v0 = (int64_t)first * pg.per_elem;
  for (int64_t w = tid; w < n; w += blockDim.x) {
    const int64_t idx = v0 + w;
    int sp = 0;
    for (int pc = 0; pc < pg.nops; ++pc) {
      const PwOp o = pg.op[pc];
      switch (o.code) {
        case PW_LOAD: {
          int64_t ob = pg.off_size == 8 ? ((const int64_t*)pg.off[o.a])[part]
                                        : (int64_t)((const int32_t*)pg.off[o.a])[part];
          stk[sp++][tid] = pw_load(pg.type[o.a], pg.base[o.a], ob + idx);
          break;
        }
        case PW_IMM: stk[sp++][tid] = pg.imm[o.a]; break;
        case PW_NEG: stk[sp - 1][tid] = pw_round(o.flags, -stk[sp - 1][tid]); break;
        case PW_STORE: {
          int64_t ob = pg.off_size == 8 ? ((const int64_t*)pg.off[o.a])[part]
                                        : (int64_t)((const int32_t*)pg.off[o.a])[part];
          pw_store(pg.type[o.a], (void*)pg.base[o.a], ob + idx, stk[--sp][tid]);
          break;
        }
        default: {
          const double y = stk[--sp][tid], x = stk[sp - 1][tid];
          double r = o.code == PW_ADD ? x + y : o.code == PW_SUB ? x - y
                   : o.code == PW_MUL ? x * y : x / y;
          stk[sp - 1][tid] = pw_round(o.flags, r);
          break;
        }
      }
    }
  }
}

// ---------------------------------------------------------------------------
// Front end
// ---------------------------------------------------------------------------

struct Fail {
  int code;
  std::string msg;
};

[[noreturn]] void fail(int code, int line, const char* fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof(buf), fmt, ap);
  va_end(ap);
  Fail f{code, line > 0 ? ("line " + std::to_string(line) + ": " + buf) : std::string(buf)};
  throw f;
}

enum TokKind { T_ID, T_NUM, T_PUNCT, T_ATTR, T_END };
struct Tok {
  TokKind k;
  std::string s;
  int line;
};

struct Macro {
  bool fn = false;
  std::vector<std::string> params;
  std::vector<Tok> body;
};

std::string strip_comments(const std::string& in) {
  std::string out;
  out.reserve(in.size());
  size_t i = 0;
  while (i < in.size()) {
    char c = in[i];
    if (c == '/' && i + 1 < in.size() && in[i + 1] == '/') {
      while (i < in.size() && in[i] != '\n') ++i;
    } else if (c == '/' && i + 1 < in.size() && in[i + 1] == '*') {
      i += 2;
      while (i + 1 < in.size() && !(in[i] == '*' && in[i + 1] == '/')) {
        if (in[i] == '\n') out += '\n';
        ++i;
      }
      i += 2;
      out += ' ';
    } else if (c == '"' || c == '\'') {
      char q = c;
      out += c;
      ++i;
      while (i < in.size() && in[i] != q) {
        if (in[i] == '\\' && i + 1 < in.size()) out += in[i++];
        out += in[i++];
      }
      if (i < in.size()) out += in[i++];
    } else {
      out += c;
      ++i;
    }
  }
  return out;
}

bool id_start(char c) { return std::isalpha((unsigned char)c) || c == '_'; }
bool id_char(char c) { return std::isalnum((unsigned char)c) || c == '_'; }

void tokenize(const std::string& s, int line0, std::vector<Tok>& out) {
  static const char* multi[] = {"<<=", ">>=", "++", "--", "+=", "-=", "*=", "/=", "%=", "<=",
                                ">=",  "==",  "!=", "&&", "||", "->", "<<", ">>"};
  int line = line0;
  size_t i = 0;
  while (i < s.size()) {
    char c = s[i];
    if (c == '\n') {
      ++line;
      ++i;
      continue;
    }
    if (std::isspace((unsigned char)c)) {
      ++i;
      continue;
    }
    if (id_start(c)) {
      size_t j = i;
      while (j < s.size() && id_char(s[j])) ++j;
      out.push_back({T_ID, s.substr(i, j - i), line});
      i = j;
      continue;
    }
    if (std::isdigit((unsigned char)c) || (c == '.' && i + 1 < s.size() && std::isdigit((unsigned char)s[i + 1]))) {
      size_t j = i;
      while (j < s.size()) {
        char d = s[j];
        if (std::isalnum((unsigned char)d) || d == '.') {
          ++j;
        } else if ((d == '+' || d == '-') && (s[j - 1] == 'e' || s[j - 1] == 'E') &&
                   !(s[i] == '0' && j > i + 1 && (s[i + 1] == 'x' || s[i + 1] == 'X'))) {
          ++j;
        } else {
          break;
        }
      }
      out.push_back({T_NUM, s.substr(i, j - i), line});
      i = j;
      continue;
    }
    if (c == '@' && i + 1 < s.size() && id_start(s[i + 1])) {
      size_t j = i + 1;
      while (j < s.size() && id_char(s[j])) ++j;
      out.push_back({T_ATTR, s.substr(i, j - i), line});
      i = j;
      continue;
    }
    bool m = false;
    for (const char* op : multi) {
      size_t L = strlen(op);
      if (s.compare(i, L, op) == 0) {
        out.push_back({T_PUNCT, op, line});
        i += L;
        m = true;
        break;
      }
    }
    if (m) continue;
    out.push_back({T_PUNCT, std::string(1, c), line});
    ++i;
  }
}

struct Prep {
  std::vector<Tok> toks;  // code tokens, object/function macros not yet expanded
  std::map<std::string, Macro> macros;
};

// Comments out, directives collected, the rest tokenized with line numbers.
Prep preprocess(const std::string& src) {
  Prep p;
  std::string s = strip_comments(src);
  std::vector<std::string> lines;
  size_t a = 0;
  while (a <= s.size()) {
    size_t b = s.find('\n', a);
    if (b == std::string::npos) b = s.size();
    lines.push_back(s.substr(a, b - a));
    a = b + 1;
  }
  std::string code;
  for (size_t li = 0; li < lines.size(); ++li) {
    std::string ln = lines[li];
    size_t f = ln.find_first_not_of(" \t\r");
    if (f == std::string::npos || ln[f] != '#') {
      code += ln;
      code += '\n';
      continue;
    }
    const int line = (int)li + 1;
    int extra = 0;
    while (!ln.empty() && (ln.back() == '\\' || ln.back() == '\r' || ln.back() == ' ')) {
      if (ln.back() == '\\') {
        ln.pop_back();
        if (li + 1 < lines.size()) {
          ln += " " + lines[++li];
          ++extra;
        }
      } else {
        ln.pop_back();
      }
    }
    code += std::string(1 + extra, '\n');
    std::string d = ln.substr(f + 1);
    size_t k = d.find_first_not_of(" \t");
    d = k == std::string::npos ? "" : d.substr(k);
    size_t e = 0;
    while (e < d.size() && id_char(d[e])) ++e;
    std::string dir = d.substr(0, e), rest = d.substr(e);
    if (dir == "pragma" || dir.empty()) continue;
    if (dir == "undef") {
      std::vector<Tok> t;
      tokenize(rest, line, t);
      if (!t.empty()) p.macros.erase(t[0].s);
      continue;
    }
    if (dir != "define") fail(DGK_ERR_UNSUPPORTED, line, "preprocessor directive #%s", dir.c_str());
    size_t q = rest.find_first_not_of(" \t");
    if (q == std::string::npos || !id_start(rest[q])) fail(DGK_ERR_INVALID, line, "bad #define");
    size_t r = q;
    while (r < rest.size() && id_char(rest[r])) ++r;
    std::string name = rest.substr(q, r - q);
    Macro mac;
    std::string body;
    if (r < rest.size() && rest[r] == '(') {
      mac.fn = true;
      size_t c = rest.find(')', r);
      if (c == std::string::npos) fail(DGK_ERR_INVALID, line, "bad macro parameters");
      std::vector<Tok> pt;
      tokenize(rest.substr(r + 1, c - r - 1), line, pt);
      for (auto& t : pt)
        if (t.k == T_ID) mac.params.push_back(t.s);
      body = rest.substr(c + 1);
    } else {
      body = rest.substr(r);
    }
    tokenize(body, line, mac.body);
    for (auto& t : mac.body)
      if (t.s == "#" || t.s == "##")
        fail(DGK_ERR_UNSUPPORTED, line, "stringizing/pasting in macro %s", name.c_str());
    p.macros[name] = mac;
  }
  tokenize(code, 1, p.toks);
  return p;
}

bool is_runtime_macro(const std::string& n) {
  return n.rfind("MATMUL_", 0) == 0 ||
         (n.size() > 5 && n.compare(n.size() - 5, 5, "_PART") == 0);
}

// Expand user macros (object-like everywhere; function-like unless they are
// the runner's X_PART / MATMUL_* forms, which the statement reader knows).
// NPART and KpartMax stay names: the loop forms are recognised by them.
void expand(const std::vector<Tok>& in, const std::map<std::string, Macro>& macros,
            std::set<std::string>& active, std::vector<Tok>& out, int depth) {
  if (depth > 32) fail(DGK_ERR_INVALID, in.empty() ? 0 : in[0].line, "macro expansion too deep");
  for (size_t i = 0; i < in.size(); ++i) {
    const Tok& t = in[i];
    auto it = t.k == T_ID ? macros.find(t.s) : macros.end();
    if (it == macros.end() || active.count(t.s) || t.s == "NPART" || t.s == "KpartMax" ||
        is_runtime_macro(t.s)) {
      out.push_back(t);
      continue;
    }
    const Macro& m = it->second;
    std::vector<Tok> body;
    if (!m.fn) {
      body = m.body;
    } else {
      if (i + 1 >= in.size() || in[i + 1].s != "(") {
        out.push_back(t);
        continue;
      }
      std::vector<std::vector<Tok>> args(1);
      int dpt = 0;
      size_t j = i + 2;
      for (; j < in.size(); ++j) {
        const std::string& s = in[j].s;
        if (dpt == 0 && s == ")") break;
        if (dpt == 0 && s == ",") {
          args.emplace_back();
          continue;
        }
        if (s == "(" || s == "[" || s == "{") ++dpt;
        if (s == ")" || s == "]" || s == "}") --dpt;
        args.back().push_back(in[j]);
      }
      if (j >= in.size()) fail(DGK_ERR_INVALID, t.line, "unterminated call of macro %s", t.s.c_str());
      if (args.size() == 1 && args[0].empty() && m.params.empty()) args.clear();
      if (args.size() != m.params.size())
        fail(DGK_ERR_INVALID, t.line, "macro %s takes %d arguments", t.s.c_str(), (int)m.params.size());
      for (const Tok& b : m.body) {
        size_t q = 0;
        for (; q < m.params.size(); ++q)
          if (b.k == T_ID && b.s == m.params[q]) break;
        if (q < m.params.size()) {
          body.push_back({T_PUNCT, "(", t.line});
          for (Tok a : args[q]) {
            a.line = t.line;
            body.push_back(a);
          }
          body.push_back({T_PUNCT, ")", t.line});
        } else {
          Tok c = b;
          c.line = t.line;
          body.push_back(c);
        }
      }
      i = j;
    }
    for (auto& b : body) b.line = t.line;
    active.insert(t.s);
    expand(body, macros, active, out, depth + 1);
    active.erase(t.s);
  }
}

// --------------------------- expressions -----------------------------------
enum CType { CT_INT = 0, CT_FLT = 1, CT_DBL = 2 };

struct Expr;
using EP = std::shared_ptr<Expr>;
struct Expr {
  enum Kind { NUM, ID, IDX, CALL, BIN, UN, CAST } k;
  std::string s;  // name / operator / cast type
  double v = 0;
  int ct = CT_DBL;  // literal type
  std::vector<EP> a;
  int line = 0;
};

struct Parser {
  const std::vector<Tok>& t;
  size_t i;
  Parser(const std::vector<Tok>& toks, size_t start) : t(toks), i(start) {}

  const Tok& peek(size_t o = 0) const {
    static const Tok end{T_END, "", 0};
    return i + o < t.size() ? t[i + o] : end;
  }
  int line() const { return peek().line ? peek().line : (t.empty() ? 0 : t.back().line); }
  bool is(const char* s, size_t o = 0) const { return peek(o).k != T_END && peek(o).s == s; }
  void expect(const char* s) {
    if (!is(s)) fail(DGK_ERR_UNSUPPORTED, line(), "expected '%s', found '%s'", s, peek().s.c_str());
    ++i;
  }
  static bool is_type(const std::string& s) {
    return s == "double" || s == "float" || s == "int" || s == "long" || s == "int_t" ||
           s == "real_t" || s == "unsigned" || s == "size_t";
  }

  EP node(Expr::Kind k, const std::string& s, int ln) {
    auto e = std::make_shared<Expr>();
    e->k = k;
    e->s = s;
    e->line = ln;
    return e;
  }
  EP bin(const std::string& op, EP l, EP r) {
    EP e = node(Expr::BIN, op, l->line);
    e->a = {l, r};
    return e;
  }

  EP primary() {
    const Tok& k = peek();
    if (k.k == T_NUM) {
      ++i;
      EP e = node(Expr::NUM, k.s, k.line);
      std::string s = k.s;
      const bool hex = s.size() > 1 && s[0] == '0' && (s[1] == 'x' || s[1] == 'X');
      bool fsuf = false;
      while (!s.empty() && strchr("fFlLuU", s.back()) && !(hex && (s.back() == 'f' || s.back() == 'F'))) {
        if (s.back() == 'f' || s.back() == 'F') fsuf = true;
        s.pop_back();
      }
      const bool real = !hex && s.find_first_of(".eE") != std::string::npos;
      e->ct = fsuf ? CT_FLT : (real ? CT_DBL : CT_INT);
      e->v = fsuf ? (double)strtof(s.c_str(), nullptr)
                  : (real ? strtod(s.c_str(), nullptr) : (double)strtoll(s.c_str(), nullptr, 0));
      return e;
    }
    if (k.k == T_ID) {
      ++i;
      return node(Expr::ID, k.s, k.line);
    }
    if (is("(")) {
      ++i;
      EP e = expr();
      expect(")");
      return e;
    }
    fail(DGK_ERR_UNSUPPORTED, k.line, "unexpected '%s' in expression", k.s.c_str());
  }
  EP postfix() {
    EP e = primary();
    for (;;) {
      if (is("[")) {
        ++i;
        EP x = expr();
        expect("]");
        EP n = node(Expr::IDX, "", e->line);
        n->a = {e, x};
        e = n;
      } else if (is("(") && e->k == Expr::ID) {
        ++i;
        EP n = node(Expr::CALL, e->s, e->line);
        if (!is(")")) {
          n->a.push_back(assign());
          while (is(",")) {
            ++i;
            n->a.push_back(assign());
          }
        }
        expect(")");
        e = n;
      } else if (is("++") || is("--")) {
        EP n = node(Expr::UN, std::string("post") + peek().s, e->line);
        ++i;
        n->a = {e};
        e = n;
      } else {
        return e;
      }
    }
  }
  EP unary() {
    if (is("-") || is("+") || is("!") || is("++") || is("--")) {
      std::string op = peek().s;
      int ln = line();
      ++i;
      EP n = node(Expr::UN, op, ln);
      n->a = {unary()};
      return n;
    }
    if (is("(") && peek(1).k == T_ID && is_type(peek(1).s)) {
      int ln = line();
      ++i;
      std::string ty = peek().s;
      ++i;
      while (peek().k == T_ID && is_type(peek().s)) ++i;
      if (is("*")) fail(DGK_ERR_UNSUPPORTED, ln, "pointer cast");
      expect(")");
      EP n = node(Expr::CAST, ty, ln);
      n->a = {unary()};
      return n;
    }
    return postfix();
  }
  EP binary(int level) {
    static const std::vector<std::vector<std::string>> ops = {
        {"||"}, {"&&"}, {"==", "!="}, {"<", "<=", ">", ">="}, {"+", "-"}, {"*", "/", "%"}};
    if (level == (int)ops.size()) return unary();
    EP l = binary(level + 1);
    for (;;) {
      bool hit = false;
      for (auto& o : ops[level])
        if (peek().k == T_PUNCT && peek().s == o) {
          ++i;
          l = bin(o, l, binary(level + 1));
          hit = true;
          break;
        }
      if (!hit) return l;
    }
  }
  EP assign() {
    EP l = binary(0);
    if (peek().k == T_PUNCT && (is("=") || is("+=") || is("-=") || is("*=") || is("/="))) {
      std::string op = peek().s;
      ++i;
      return bin(op, l, assign());
    }
    if (is("?")) fail(DGK_ERR_UNSUPPORTED, line(), "conditional expression");
    return l;
  }
  EP expr() { return assign(); }
};

// ---------------------------- statements -----------------------------------
struct Stmt;
using SP = std::shared_ptr<Stmt>;
struct Stmt {
  enum Kind { BLOCK, FOR, IF, DECL, EXPR, EMPTY } k;
  int line = 0;
  std::vector<SP> body;  // BLOCK: statements; FOR/IF: body[0]
  // FOR
  std::string var, attr;
  EP init, cond, incr;
  // DECL
  bool is_const = false, ptr = false;
  std::string type, name;
  // DECL init / EXPR / IF cond
  EP e;
};

struct StmtParser : Parser {
  using Parser::Parser;

  SP block() {
    auto b = std::make_shared<Stmt>();
    b->k = Stmt::BLOCK;
    b->line = line();
    expect("{");
    while (!is("}")) {
      if (peek().k == T_END) fail(DGK_ERR_INVALID, b->line, "unterminated block");
      b->body.push_back(stmt());
    }
    ++i;
    return b;
  }
  bool at_decl() const {
    size_t o = 0;
    if (is("const", o)) ++o;
    return peek(o).k == T_ID && is_type(peek(o).s);
  }
  SP decl(bool need_semi) {
    auto d = std::make_shared<Stmt>();
    d->k = Stmt::DECL;
    d->line = line();
    if (is("const")) {
      d->is_const = true;
      ++i;
    }
    d->type = peek().s;
    ++i;
    while (peek().k == T_ID && is_type(peek().s)) d->type += " " + t[i++].s;
    if (is("const")) ++i;
    if (is("*")) {
      d->ptr = true;
      ++i;
      if (is("const")) ++i;
      if (is("__restrict__") || is("restrict")) ++i;
    }
    if (peek().k != T_ID) fail(DGK_ERR_UNSUPPORTED, d->line, "declaration without a name");
    d->name = t[i++].s;
    if (is("[")) fail(DGK_ERR_UNSUPPORTED, d->line, "local array %s", d->name.c_str());
    if (is("=")) {
      ++i;
      d->e = expr();
    }
    if (is(",")) fail(DGK_ERR_UNSUPPORTED, d->line, "several declarators in one declaration");
    if (need_semi) expect(";");
    return d;
  }
  SP stmt() {
    int ln = line();
    if (is("{")) return block();
    if (is(";")) {
      ++i;
      auto s = std::make_shared<Stmt>();
      s->k = Stmt::EMPTY;
      s->line = ln;
      return s;
    }
    if (is("for")) {
      ++i;
      auto f = std::make_shared<Stmt>();
      f->k = Stmt::FOR;
      f->line = ln;
      expect("(");
      if (at_decl()) {
        SP d = decl(false);
        if (d->ptr || !d->e) fail(DGK_ERR_UNSUPPORTED, ln, "loop variable must be an integer");
        f->var = d->name;
        f->init = d->e;
      } else {
        EP e = expr();
        if (e->k != Expr::BIN || e->s != "=" || e->a[0]->k != Expr::ID)
          fail(DGK_ERR_UNSUPPORTED, ln, "loop initialiser");
        f->var = e->a[0]->s;
        f->init = e->a[1];
      }
      expect(";");
      f->cond = expr();
      expect(";");
      f->incr = expr();
      if (is(";")) {
        ++i;
        if (peek().k != T_ATTR) fail(DGK_ERR_UNSUPPORTED, ln, "expected a loop attribute");
        f->attr = t[i++].s;
      }
      expect(")");
      f->body.push_back(stmt());
      return f;
    }
    if (is("if")) {
      ++i;
      auto s = std::make_shared<Stmt>();
      s->k = Stmt::IF;
      s->line = ln;
      expect("(");
      s->e = expr();
      expect(")");
      s->body.push_back(stmt());
      if (is("else")) fail(DGK_ERR_UNSUPPORTED, ln, "else branch");
      return s;
    }
    if (is("while") || is("do") || is("return") || is("break") || is("continue") || is("switch"))
      fail(DGK_ERR_UNSUPPORTED, ln, "'%s' statement", peek().s.c_str());
    if (at_decl()) return decl(true);
    auto s = std::make_shared<Stmt>();
    s->k = Stmt::EXPR;
    s->line = ln;
    s->e = expr();
    expect(";");
    return s;
  }
};

// ------------------------------ model ---------------------------------------
struct Param {
  bool is_const = false, ptr = false;
  std::string type, name;
};

struct Array {
  std::string name;
  int gparam = -1, oparam = -1;
  int type = ET_F64;
  bool is_const = false;
};

struct Matrix {
  std::string name;
  bool is_static = false;
  int param = -1;         // device matrix: signature position
  int rows = 0, cols = 0;
  int type = ET_F64;
  std::vector<double> colmajor;  // static values, A(i,j) at [j*rows + i]
  void* dev = nullptr;           // static: device copy (column-major, own type)
};

struct MacroInfo {  // MATMUL_X / MATMUL_ADD_X of the preamble
  std::string matrix;
  bool add = false, is_static = false;
  int rows = 0, cols = 0, atype = ET_F64;
};

enum OpKind { OP_MATMUL, OP_PW, OP_DERIV };
struct Op {
  OpKind k;
  int line = 0;
  // OP_MATMUL
  int mat = -1, in = -1, out = -1, rs = 0, cs = 0, atype = ET_F64;
  bool add = false;
  // OP_PW
  int per_elem = 0;
  std::vector<PwOp> code;
  std::vector<double> imm;
  std::vector<int> imm_param;  // per imm slot: scalar parameter position, or -1
  std::vector<int> arr;        // per array slot: array index
  std::string text;            // description
  // OP_DERIV
  int dmat[3] = {-1, -1, -1}, u = -1, temps[3] = {-1, -1, -1}, mets[9] = {}, outs[3] = {};
};

int elem_type_of(const std::string& ty, int int_size, int line) {
  if (ty == "double" || ty == "real_t") return ET_F64;
  if (ty == "float") return ET_F32;
  if (ty == "int") return ET_I32;
  if (ty == "long" || ty == "long long") return ET_I64;
  if (ty == "int_t") return int_size == 8 ? ET_I64 : ET_I32;
  fail(DGK_ERR_UNSUPPORTED, line, "type '%s'", ty.c_str());
}

const char* et_name(int t) {
  static const char* n[] = {"double", "float", "int", "long"};
  return n[t & 3];
}

int np_order(int np) {
  for (int N = 1; N <= 6; ++N)
    if ((N + 1) * (N + 2) * (N + 3) / 6 == np) return N;
  return -1;
}

}  // namespace

struct dgk_kernel_s {
  dgk_device dev = nullptr;
  std::string name;
  int int_size = 8, npart = 0, kpart_max = 0;
  std::vector<Param> params;
  std::vector<Array> arrays;
  std::vector<Matrix> mats;
  std::vector<Op> ops;
};

namespace {

struct Builder {
  dgk_kernel_s& k;
  std::map<std::string, Macro> macros;
  std::map<std::string, MacroInfo> mm;          // MATMUL macro name -> info
  std::map<std::string, int> part_type;         // X of X_PART -> element type
  std::map<std::string, std::vector<double>> static_vals;  // name -> column-major values
  std::map<std::string, std::pair<int, int>> static_dims;  // name -> (cols, rows) as declared
  std::map<std::string, int> static_type;
  std::map<std::string, int> alias;             // alias -> array index
  std::string part_var;
  bool have_int_t = false;

  explicit Builder(dgk_kernel_s& kk) : k(kk) {}

  // ---- constants
  bool const_eval(const EP& e, long long& v) const {
    switch (e->k) {
      case Expr::NUM:
        if (e->ct != CT_INT) return false;
        v = (long long)e->v;
        return true;
      case Expr::ID:
        if (e->s == "NPART") { v = k.npart; return true; }
        if (e->s == "KpartMax") { v = k.kpart_max; return true; }
        return false;
      case Expr::UN: {
        long long x;
        if (e->s != "-" || !const_eval(e->a[0], x)) return false;
        v = -x;
        return true;
      }
      case Expr::BIN: {
        long long x, y;
        if (!const_eval(e->a[0], x) || !const_eval(e->a[1], y)) return false;
        if (e->s == "+") v = x + y;
        else if (e->s == "-") v = x - y;
        else if (e->s == "*") v = x * y;
        else if (e->s == "/" && y != 0) v = x / y;
        else return false;
        return true;
      }
      default:
        return false;
    }
  }
  bool is_id(const EP& e, const std::string& n) const { return e && e->k == Expr::ID && e->s == n; }
  bool is_zero(const EP& e) const {
    long long v;
    return e->k == Expr::NUM ? e->v == 0.0 : (const_eval(e, v) && v == 0);
  }
  // K[part]
  bool is_kpart(const EP& e) const {
    return e->k == Expr::IDX && is_id(e->a[0], "K") && is_id(e->a[1], part_var);
  }
  // var < bound with var starting at 0 and stepping by one
  void loop_shape(const Stmt& f, EP& bound) const {
    if (!is_zero(f.init)) fail(DGK_ERR_UNSUPPORTED, f.line, "loop %s must start at 0", f.var.c_str());
    const EP& c = f.cond;
    if (c->k != Expr::BIN || c->s != "<" || !is_id(c->a[0], f.var))
      fail(DGK_ERR_UNSUPPORTED, f.line, "loop condition must be %s < bound", f.var.c_str());
    const EP& s = f.incr;
    bool ok = (s->k == Expr::UN && (s->s == "++" || s->s == "post++") && is_id(s->a[0], f.var)) ||
              (s->k == Expr::BIN && s->s == "+=" && is_id(s->a[0], f.var) && s->a[1]->k == Expr::NUM &&
               s->a[1]->v == 1.0);
    if (!ok) fail(DGK_ERR_UNSUPPORTED, f.line, "loop %s must step by one", f.var.c_str());
    bound = c->a[1];
  }
  // a*b form in either order, with a == var and b a constant C
  bool is_var_times_const(const EP& e, const std::string& var, long long& C) const {
    if (e->k != Expr::BIN || e->s != "*") return false;
    if (is_id(e->a[0], var) && const_eval(e->a[1], C)) return true;
    if (is_id(e->a[1], var) && const_eval(e->a[0], C)) return true;
    return false;
  }
  // x + y*C or y*C + x (x, y variable names)
  bool is_lin(const EP& e, const std::string& x, const std::string& y, long long C) const {
    if (!e || e->k != Expr::BIN || e->s != "+") return false;
    long long c;
    if (is_id(e->a[0], x) && is_var_times_const(e->a[1], y, c) && c == C) return true;
    if (is_id(e->a[1], x) && is_var_times_const(e->a[0], y, c) && c == C) return true;
    return false;
  }
  static std::vector<SP> stmts_of(const SP& s) {
    if (s->k == Stmt::BLOCK) {
      std::vector<SP> out;
      for (auto& x : s->body)
        if (x->k != Stmt::EMPTY) out.push_back(x);
      return out;
    }
    return {s};
  }

  int alias_of(const EP& e, int line) const {
    if (e->k != Expr::ID) fail(DGK_ERR_UNSUPPORTED, line, "expected an array alias");
    auto it = alias.find(e->s);
    if (it == alias.end())
      fail(DGK_ERR_UNSUPPORTED, line, "'%s' is not an array alias declared with X_PART(part)",
           e->s.c_str());
    return it->second;
  }
  int param_pos(const std::string& n) const {
    for (size_t i = 0; i < k.params.size(); ++i)
      if (k.params[i].name == n) return (int)i;
    return -1;
  }
  int matrix_index(const std::string& n, int line) {
    for (size_t i = 0; i < k.mats.size(); ++i)
      if (k.mats[i].name == n) return (int)i;
    Matrix m;
    m.name = n;
    int p = param_pos(n);
    if (p >= 0) {
      if (!k.params[p].ptr) fail(DGK_ERR_INVALID, line, "matrix %s is not a pointer argument", n.c_str());
      m.param = p;
      m.type = elem_type_of(k.params[p].type, k.int_size, line);
      if (m.type != ET_F64 && m.type != ET_F32)
        fail(DGK_ERR_UNSUPPORTED, line, "matrix %s of type %s", n.c_str(), k.params[p].type.c_str());
    } else if (static_vals.count(n)) {
      m.is_static = true;
      m.type = static_type[n];
      m.cols = static_dims[n].first;
      m.rows = static_dims[n].second;
      m.colmajor = static_vals[n];
    } else {
      fail(DGK_ERR_INVALID, line, "matrix %s is neither an argument nor a static matrix", n.c_str());
    }
    k.mats.push_back(m);
    return (int)k.mats.size() - 1;
  }

  // ---- pointwise compiler
  struct PwBuild {
    Op op;
    std::map<int, int> arr_slot;
    int depth = 0, max_depth = 0;
  };
  int pw_arr_slot(PwBuild& b, int arr, int line) {
    auto it = b.arr_slot.find(arr);
    if (it != b.arr_slot.end()) return it->second;
    if ((int)b.op.arr.size() >= kPwMaxArr)
      fail(DGK_ERR_UNSUPPORTED, line, "more than %d arrays in one loop", kPwMaxArr);
    b.op.arr.push_back(arr);
    return b.arr_slot[arr] = (int)b.op.arr.size() - 1;
  }
  void emit(PwBuild& b, uint8_t code, uint8_t flags, int a, int line) {
    if ((int)b.op.code.size() >= kPwMaxOps)
      fail(DGK_ERR_UNSUPPORTED, line, "loop body longer than %d operations", kPwMaxOps);
    b.op.code.push_back(PwOp{code, flags, (uint16_t)a});
    if (code == PW_LOAD || code == PW_IMM) ++b.depth;
    else if (code == PW_STORE || (code != PW_NEG)) --b.depth;
    if (b.depth > b.max_depth) b.max_depth = b.depth;
    if (b.max_depth > kPwMaxDepth) fail(DGK_ERR_UNSUPPORTED, line, "expression too deep");
  }
  int imm_slot(PwBuild& b, double v, int param, int line) {
    for (size_t i = 0; i < b.op.imm.size(); ++i)
      if (b.op.imm_param[i] == param && (param >= 0 || b.op.imm[i] == v)) return (int)i;
    if ((int)b.op.imm.size() >= kPwMaxImm) fail(DGK_ERR_UNSUPPORTED, line, "too many constants");
    b.op.imm.push_back(v);
    b.op.imm_param.push_back(param);
    return (int)b.op.imm.size() - 1;
  }
  static int ct_of_et(int et) { return et == ET_F64 ? CT_DBL : et == ET_F32 ? CT_FLT : CT_INT; }
  // index forms accepted for node reads/writes
  std::function<bool(const EP&)> is_node_index;

  int pw_expr(PwBuild& b, const EP& e) {
    switch (e->k) {
      case Expr::NUM:
        emit(b, PW_IMM, 0, imm_slot(b, e->v, -1, e->line), e->line);
        return e->ct;
      case Expr::ID: {
        long long c;
        if (const_eval(e, c)) {
          emit(b, PW_IMM, 0, imm_slot(b, (double)c, -1, e->line), e->line);
          return CT_INT;
        }
        int p = param_pos(e->s);
        if (p < 0 || k.params[p].ptr)
          fail(DGK_ERR_UNSUPPORTED, e->line, "'%s' is not a scalar argument", e->s.c_str());
        emit(b, PW_IMM, 0, imm_slot(b, 0.0, p, e->line), e->line);
        return ct_of_et(elem_type_of(k.params[p].type, k.int_size, e->line));
      }
      case Expr::IDX: {
        if (!is_node_index(e->a[1]))
          fail(DGK_ERR_UNSUPPORTED, e->line, "array read away from the loop's node index");
        int arr = alias_of(e->a[0], e->line);
        emit(b, PW_LOAD, 0, pw_arr_slot(b, arr, e->line), e->line);
        return ct_of_et(k.arrays[arr].type);
      }
      case Expr::UN: {
        if (e->s == "+") return pw_expr(b, e->a[0]);
        if (e->s != "-") fail(DGK_ERR_UNSUPPORTED, e->line, "operator %s", e->s.c_str());
        int t = pw_expr(b, e->a[0]);
        emit(b, PW_NEG, t == CT_FLT ? PWF_F32 : 0, 0, e->line);
        return t;
      }
      case Expr::CAST: {
        int t = pw_expr(b, e->a[0]);
        int to = ct_of_et(elem_type_of(e->s, k.int_size, e->line));
        if (to == CT_FLT && t != CT_FLT) {  // (float)x: round to FP32
          emit(b, PW_IMM, 0, imm_slot(b, 0.0, -1, e->line), e->line);
          emit(b, PW_ADD, PWF_F32, 0, e->line);
        } else if (to == CT_INT && t != CT_INT) {
          fail(DGK_ERR_UNSUPPORTED, e->line, "cast to an integer type");
        }
        return to;
      }
      case Expr::BIN: {
        uint8_t code = e->s == "+" ? PW_ADD : e->s == "-" ? PW_SUB : e->s == "*" ? PW_MUL
                     : e->s == "/" ? PW_DIV : 255;
        if (code == 255) fail(DGK_ERR_UNSUPPORTED, e->line, "operator %s", e->s.c_str());
        int ta = pw_expr(b, e->a[0]);
        int tb = pw_expr(b, e->a[1]);
        int t = std::max(ta, tb);
        uint8_t f = t == CT_FLT ? PWF_F32 : (t == CT_INT && code == PW_DIV ? PWF_INT : 0);
        emit(b, code, f, 0, e->line);
        return t;
      }
      default:
        fail(DGK_ERR_UNSUPPORTED, e->line, "expression form");
    }
  }
  void pw_assign(PwBuild& b, const SP& s) {
    if (s->k != Stmt::EXPR || s->e->k != Expr::BIN)
      fail(DGK_ERR_UNSUPPORTED, s->line, "loop body statement must be an assignment");
    const EP& e = s->e;
    const std::string& op = e->s;
    if (op != "=" && op != "+=" && op != "-=" && op != "*=" && op != "/=")
      fail(DGK_ERR_UNSUPPORTED, s->line, "loop body statement must be an assignment");
    const EP& lhs = e->a[0];
    if (lhs->k != Expr::IDX || !is_node_index(lhs->a[1]))
      fail(DGK_ERR_UNSUPPORTED, s->line, "assignment target must be A[node index]");
    int arr = alias_of(lhs->a[0], s->line);
    if (k.arrays[arr].is_const)
      fail(DGK_ERR_INVALID, s->line, "assignment to const array %s", k.arrays[arr].name.c_str());
    int slot = pw_arr_slot(b, arr, s->line);
    int tl = ct_of_et(k.arrays[arr].type);
    if (op == "=") {
      pw_expr(b, e->a[1]);
    } else {
      emit(b, PW_LOAD, 0, slot, s->line);
      int tr = pw_expr(b, e->a[1]);
      int t = std::max(tl, tr);
      uint8_t code = op == "+=" ? PW_ADD : op == "-=" ? PW_SUB : op == "*=" ? PW_MUL : PW_DIV;
      emit(b, code, t == CT_FLT ? PWF_F32 : (t == CT_INT && code == PW_DIV ? PWF_INT : 0), 0, s->line);
    }
    emit(b, PW_STORE, 0, slot, s->line);
    b.op.text += (b.op.text.empty() ? "" : "; ") + k.arrays[arr].name + " " + op + " ...";
  }

  // ---- statements in the partition loop
  void read_part_body(const std::vector<SP>& S) {
    for (const SP& s : S) {
      if (s->k == Stmt::EMPTY) continue;
      if (s->k == Stmt::DECL) {
        const EP& in = s->e;
        if (!s->ptr || !in || in->k != Expr::CALL || in->a.size() != 1 || !is_id(in->a[0], part_var) ||
            in->s.size() <= 5 || in->s.compare(in->s.size() - 5, 5, "_PART") != 0)
          fail(DGK_ERR_UNSUPPORTED, s->line, "declaration '%s' (only T* A = X_PART(part) is known)",
               s->name.c_str());
        std::string X = in->s.substr(0, in->s.size() - 5);
        int ai = -1;
        for (size_t i = 0; i < k.arrays.size(); ++i)
          if (k.arrays[i].name == X) ai = (int)i;
        if (ai < 0) fail(DGK_ERR_INVALID, s->line, "array %s is not in the kernel signature", X.c_str());
        int dt = elem_type_of(s->type, k.int_size, s->line);
        if (dt != k.arrays[ai].type)
          fail(DGK_ERR_INVALID, s->line, "%s declared %s but the array holds %s", s->name.c_str(),
               s->type.c_str(), et_name(k.arrays[ai].type));
        if (!s->is_const && k.arrays[ai].is_const)
          fail(DGK_ERR_INVALID, s->line, "non-const alias %s of const array %s", s->name.c_str(), X.c_str());
        alias[s->name] = ai;
        continue;
      }
      if (s->k == Stmt::EXPR && s->e->k == Expr::CALL && s->e->s.rfind("MATMUL_", 0) == 0) {
        read_matmul(s);
        continue;
      }
      if (s->k == Stmt::FOR) {
        read_loop(s);
        continue;
      }
      fail(DGK_ERR_UNSUPPORTED, s->line, "statement in the partition loop");
    }
  }

  void read_matmul(const SP& s) {
    const EP& c = s->e;
    auto it = mm.find(c->s);
    if (it == mm.end()) fail(DGK_ERR_INVALID, s->line, "no macro %s in the preamble", c->s.c_str());
    if (c->a.size() != 3 || !is_kpart(c->a[2]))
      fail(DGK_ERR_UNSUPPORTED, s->line, "%s(IN, OUT, K[part]) expected", c->s.c_str());
    const MacroInfo& mi = it->second;
    Op op;
    op.k = OP_MATMUL;
    op.line = s->line;
    op.mat = matrix_index(mi.matrix, s->line);
    Matrix& M = k.mats[op.mat];
    if (M.is_static != mi.is_static)
      fail(DGK_ERR_INVALID, s->line, "macro %s and matrix %s disagree on static/device", c->s.c_str(),
           M.name.c_str());
    if (M.is_static && (M.rows != mi.rows || M.cols != mi.cols))
      fail(DGK_ERR_INVALID, s->line, "macro %s is %dx%d, static matrix %s is %dx%d", c->s.c_str(),
           mi.rows, mi.cols, M.name.c_str(), M.rows, M.cols);
    M.rows = mi.rows;
    M.cols = mi.cols;
    op.in = alias_of(c->a[0], s->line);
    op.out = alias_of(c->a[1], s->line);
    if (k.arrays[op.out].is_const)
      fail(DGK_ERR_INVALID, s->line, "MATMUL into const array %s", k.arrays[op.out].name.c_str());
    for (int x : {op.in, op.out})
      if (k.arrays[x].type != ET_F64 && k.arrays[x].type != ET_F32)
        fail(DGK_ERR_UNSUPPORTED, s->line, "MATMUL on integer array %s", k.arrays[x].name.c_str());
    op.add = mi.add;
    op.rs = 1;
    op.cs = M.rows;
    op.atype = mi.atype;
    op.text = c->s;
    k.ops.push_back(op);
  }

  void read_loop(const SP& f) {
    EP bound;
    loop_shape(*f, bound);
    long long C = 0;
    // element matvec: for (e < K[part]) for (i < C; @inner) { ... }
    if (f->attr.empty() && is_kpart(bound)) {
      read_matvec(f);
      return;
    }
    if (f->attr != "@inner")
      fail(DGK_ERR_UNSUPPORTED, f->line, "loop over %s (known: element-node, flat-node and element "
           "matvec loops)", f->var.c_str());
    if (is_id(bound, "KpartMax")) {
      // element-node loop
      std::vector<SP> b1 = stmts_of(f->body[0]);
      if (b1.size() != 1 || b1[0]->k != Stmt::IF)
        fail(DGK_ERR_UNSUPPORTED, f->line, "element loop body must be if (%s < K[part]) {...}", f->var.c_str());
      const SP& g = b1[0];
      if (g->e->k != Expr::BIN || g->e->s != "<" || !is_id(g->e->a[0], f->var) || !is_kpart(g->e->a[1]))
        fail(DGK_ERR_UNSUPPORTED, g->line, "guard must be %s < K[part]", f->var.c_str());
      std::vector<SP> b2 = stmts_of(g->body[0]);
      if (b2.size() != 1 || b2[0]->k != Stmt::FOR || !b2[0]->attr.empty())
        fail(DGK_ERR_UNSUPPORTED, g->line, "expected a node loop inside the element guard");
      const SP& nl = b2[0];
      EP nb;
      loop_shape(*nl, nb);
      if (!const_eval(nb, C) || C <= 0) fail(DGK_ERR_UNSUPPORTED, nl->line, "node loop bound must be a constant");
      std::vector<SP> body = stmts_of(nl->body[0]);
      std::string ivar;
      size_t first = 0;
      if (!body.empty() && body[0]->k == Stmt::DECL && !body[0]->ptr && body[0]->e &&
          is_lin(body[0]->e, nl->var, f->var, C)) {
        ivar = body[0]->name;
        first = 1;
      }
      const std::string n = nl->var, e = f->var;
      is_node_index = [this, ivar, n, e, C](const EP& x) {
        return (!ivar.empty() && is_id(x, ivar)) || is_lin(x, n, e, C);
      };
      compile_pw(std::vector<SP>(body.begin() + first, body.end()), (int)C, f->line, "element-node");
      return;
    }
    long long c1;
    bool flat = bound->k == Expr::BIN && bound->s == "*" &&
                ((is_id(bound->a[1], "KpartMax") && const_eval(bound->a[0], c1)) ||
                 (is_id(bound->a[0], "KpartMax") && const_eval(bound->a[1], c1)));
    if (!flat) fail(DGK_ERR_UNSUPPORTED, f->line, "@inner loop bound (known: KpartMax, C*KpartMax)");
    const std::string v = f->var;
    is_node_index = [this, v](const EP& x) { return is_id(x, v); };
    compile_pw(stmts_of(f->body[0]), (int)c1, f->line, "flat-node");
  }

  void compile_pw(const std::vector<SP>& body, int C, int line, const char* form) {
    if (body.empty()) fail(DGK_ERR_UNSUPPORTED, line, "empty loop body");
    PwBuild b;
    b.op.k = OP_PW;
    b.op.line = line;
    b.op.per_elem = C;
    for (const SP& s : body) pw_assign(b, s);
    b.op.text = std::string(form) + " loop x" + std::to_string(C) + ": " + b.op.text;
    k.ops.push_back(b.op);
  }

  void read_matvec(const SP& f) {
    std::vector<SP> b1 = stmts_of(f->body[0]);
    if (b1.size() != 1 || b1[0]->k != Stmt::FOR || b1[0]->attr != "@inner")
      fail(DGK_ERR_UNSUPPORTED, f->line, "element loop body must be one @inner row loop");
    const SP& rl = b1[0];
    EP rb;
    loop_shape(*rl, rb);
    long long R;
    if (!const_eval(rb, R) || R <= 0) fail(DGK_ERR_UNSUPPORTED, rl->line, "row loop bound must be a constant");
    std::vector<SP> b2 = stmts_of(rl->body[0]);
    if (b2.size() != 3 || b2[0]->k != Stmt::DECL || b2[0]->ptr || !b2[0]->e || !is_zero(b2[0]->e) ||
        b2[1]->k != Stmt::FOR || b2[2]->k != Stmt::EXPR)
      fail(DGK_ERR_UNSUPPORTED, rl->line, "row body must be: T sum = 0; for (j ...) sum += ...; OUT[...] = sum;");
    const std::string sum = b2[0]->name;
    int atype = elem_type_of(b2[0]->type, k.int_size, b2[0]->line);
    if (atype != ET_F64 && atype != ET_F32) fail(DGK_ERR_UNSUPPORTED, b2[0]->line, "integer accumulator");
    const SP& jl = b2[1];
    EP jb;
    loop_shape(*jl, jb);
    long long Cc;
    if (!jl->attr.empty() || !const_eval(jb, Cc) || Cc <= 0)
      fail(DGK_ERR_UNSUPPORTED, jl->line, "column loop bound must be a constant");
    std::vector<SP> b3 = stmts_of(jl->body[0]);
    if (b3.size() != 1 || b3[0]->k != Stmt::EXPR)
      fail(DGK_ERR_UNSUPPORTED, jl->line, "column loop body must be sum += M[...] * IN[...]");
    const EP& acc = b3[0]->e;
    if (acc->k != Expr::BIN || acc->s != "+=" || !is_id(acc->a[0], sum) || acc->a[1]->k != Expr::BIN ||
        acc->a[1]->s != "*")
      fail(DGK_ERR_UNSUPPORTED, b3[0]->line, "column loop body must be %s += M[...] * IN[...]", sum.c_str());
    EP pm = acc->a[1]->a[0], pu = acc->a[1]->a[1];
    if (pm->k == Expr::IDX && alias.count(pm->a[0]->s)) std::swap(pm, pu);
    if (pm->k != Expr::IDX || pu->k != Expr::IDX || pm->a[0]->k != Expr::ID)
      fail(DGK_ERR_UNSUPPORTED, b3[0]->line, "expected M[...] * IN[...]");
    const std::string e = f->var, i = rl->var, j = jl->var;
    Op op;
    op.k = OP_MATMUL;
    op.line = f->line;
    op.mat = matrix_index(pm->a[0]->s, b3[0]->line);
    Matrix& M = k.mats[op.mat];
    if (M.is_static) fail(DGK_ERR_UNSUPPORTED, b3[0]->line, "static matrix %s indexed as a vector", M.name.c_str());
    if (is_lin(pm->a[1], j, i, Cc)) {  // M[i*C + j]
      op.rs = (int)Cc;
      op.cs = 1;
    } else if (is_lin(pm->a[1], i, j, R)) {  // M[j*R + i]
      op.rs = 1;
      op.cs = (int)R;
    } else {
      fail(DGK_ERR_UNSUPPORTED, b3[0]->line, "matrix index must be %s*%lld + %s or %s*%lld + %s",
           i.c_str(), Cc, j.c_str(), j.c_str(), R, i.c_str());
    }
    if ((M.rows && M.rows != R) || (M.cols && M.cols != Cc))
      fail(DGK_ERR_INVALID, b3[0]->line, "matrix %s used as %lldx%lld and %dx%d", M.name.c_str(), R, Cc,
           M.rows, M.cols);
    M.rows = (int)R;
    M.cols = (int)Cc;
    if (!is_lin(pu->a[1], j, e, Cc))
      fail(DGK_ERR_UNSUPPORTED, b3[0]->line, "input index must be %s*%lld + %s", e.c_str(), Cc, j.c_str());
    op.in = alias_of(pu->a[0], b3[0]->line);
    const EP& st = b2[2]->e;
    if (st->k != Expr::BIN || (st->s != "=" && st->s != "+=") || !is_id(st->a[1], sum) ||
        st->a[0]->k != Expr::IDX || !is_lin(st->a[0]->a[1], i, e, R))
      fail(DGK_ERR_UNSUPPORTED, b2[2]->line, "row result must be OUT[%s*%lld + %s] = %s", e.c_str(), R,
           i.c_str(), sum.c_str());
    op.out = alias_of(st->a[0]->a[0], b2[2]->line);
    if (k.arrays[op.out].is_const) fail(DGK_ERR_INVALID, b2[2]->line, "store into const array");
    op.add = st->s == "+=";
    op.atype = atype;
    op.text = "element matvec " + M.name;
    k.ops.push_back(op);
  }

  // ---- the runner's derivative body -> one nodal-derivative launch
  void fuse() {
    std::vector<Op> out;
    for (size_t x = 0; x < k.ops.size(); ++x) {
      if (x + 3 < k.ops.size() && try_fuse(x, out)) {
        x += 3;
        continue;
      }
      out.push_back(k.ops[x]);
    }
    k.ops.swap(out);
  }
  bool try_fuse(size_t x, std::vector<Op>& out) {
    const Op* m[3] = {&k.ops[x], &k.ops[x + 1], &k.ops[x + 2]};
    const Op& pw = k.ops[x + 3];
    for (auto* o : m) {
      if (o->k != OP_MATMUL || o->add || o->atype != ET_F64 || o->in != m[0]->in) return false;
      const Matrix& M = k.mats[o->mat];
      if (M.type != ET_F64 || M.rows != M.cols || np_order(M.rows) < 0 || o->rs != 1 || o->cs != M.rows)
        return false;
      if (k.arrays[o->out].type != ET_F64) return false;
    }
    const int np = k.mats[m[0]->mat].rows;
    if (k.mats[m[1]->mat].rows != np || k.mats[m[2]->mat].rows != np) return false;
    if (k.arrays[m[0]->in].type != ET_F64) return false;
    int T[3] = {m[0]->out, m[1]->out, m[2]->out};
    if (T[0] == T[1] || T[1] == T[2] || T[0] == T[2] || T[0] == m[0]->in || T[1] == m[0]->in ||
        T[2] == m[0]->in)
      return false;
    if (pw.k != OP_PW || pw.per_elem != np || pw.code.size() != 36 || !pw.imm.empty()) return false;
    Op d;
    d.k = OP_DERIV;
    d.line = m[0]->line;
    d.u = m[0]->in;
    for (int q = 0; q < 3; ++q) {
      d.dmat[q] = m[q]->mat;
      d.temps[q] = T[q];
    }
    int mt = -1, ot = -1;
    std::set<int> seen(T, T + 3);
    seen.insert(d.u);
    for (int r = 0; r < 3; ++r) {
      const PwOp* c = &pw.code[12 * r];
      const uint8_t want[12] = {PW_LOAD, PW_LOAD, PW_MUL, PW_LOAD, PW_LOAD, PW_MUL,
                                PW_ADD,  PW_LOAD, PW_LOAD, PW_MUL, PW_ADD, PW_STORE};
      for (int q = 0; q < 12; ++q)
        if (c[q].code != want[q] || c[q].flags) return false;
      for (int q = 0; q < 3; ++q) {
        int met = pw.arr[c[q == 0 ? 0 : (q == 1 ? 3 : 7)].a];
        int tmp = pw.arr[c[q == 0 ? 1 : (q == 1 ? 4 : 8)].a];
        if (tmp != T[q]) return false;
        if (seen.count(met)) return false;
        int t = k.arrays[met].type;
        if ((t != ET_F64 && t != ET_F32) || (mt >= 0 && t != mt)) return false;
        mt = t;
        d.mets[3 * q + r] = met;  // rx ry rz | sx sy sz | tx ty tz
      }
      int o = pw.arr[c[11].a];
      int t = k.arrays[o].type;
      if ((t != ET_F64 && t != ET_F32) || (ot >= 0 && t != ot)) return false;
      ot = t;
      d.outs[r] = o;
    }
    std::set<int> mets(d.mets, d.mets + 9), outs(d.outs, d.outs + 3);
    if (mets.size() != 9 || outs.size() != 3) return false;
    for (int o : outs)
      if (mets.count(o) || seen.count(o)) return false;
    d.text = "physical derivative (" + k.mats[d.dmat[0]].name + ", " + k.mats[d.dmat[1]].name + ", " +
             k.mats[d.dmat[2]].name + ") of " + k.arrays[d.u].name + " into " +
             k.arrays[d.outs[0]].name + ", " + k.arrays[d.outs[1]].name + ", " +
             k.arrays[d.outs[2]].name + "; temporaries " + k.arrays[T[0]].name + ", " +
             k.arrays[T[1]].name + ", " + k.arrays[T[2]].name;
    out.push_back(d);
    return true;
  }

  // ---- top level
  void read_macro_info() {
    for (auto& kv : macros) {
      const std::string& n = kv.first;
      const Macro& m = kv.second;
      if (n.size() > 5 && n.compare(n.size() - 5, 5, "_PART") == 0 && m.fn) {
        // ((T*)(X_global + X_offsets[part]))
        const auto& b = m.body;
        for (size_t i = 0; i + 2 < b.size(); ++i)
          if (b[i].s == "(" && b[i + 1].k == T_ID && Parser::is_type(b[i + 1].s) && b[i + 2].s == "*") {
            part_type[n.substr(0, n.size() - 5)] = elem_type_of(b[i + 1].s, k.int_size, b[i].line);
            break;
          }
        continue;
      }
      if (n.rfind("MATMUL_", 0) != 0 || !m.fn) continue;
      MacroInfo mi;
      const auto& b = m.body;
      for (size_t i = 0; i + 4 < b.size(); ++i) {
        if (b[i].s == "ii" && b[i + 1].s == "<" && b[i + 2].k == T_NUM) mi.rows = atoi(b[i + 2].s.c_str());
        if (b[i].s == "jj" && b[i + 1].s == "<" && b[i + 2].k == T_NUM) mi.cols = atoi(b[i + 2].s.c_str());
        if (b[i].k == T_ID && Parser::is_type(b[i].s) && b[i + 1].s == "sum" && b[i + 2].s == "=")
          mi.atype = elem_type_of(b[i].s, k.int_size, b[i].line);
        if (b[i].k == T_ID && b[i + 1].s == "[" && b[i + 2].s == "jj" && b[i + 3].s == "]" &&
            b[i + 4].s == "[") {
          mi.matrix = b[i].s;
          mi.is_static = true;
        }
        if (b[i].k == T_ID && b[i + 1].s == "[" && b[i + 2].s == "jj" && b[i + 3].s == "*") {
          mi.matrix = b[i].s;
          mi.is_static = false;
        }
        if (b[i].s == "+=" && b[i + 1].s == "sum") mi.add = true;
      }
      if (mi.matrix.empty() || mi.rows <= 0 || mi.cols <= 0) continue;  // not the runner's form
      mm[n] = mi;
    }
  }

  void read_top(const std::vector<Tok>& T, const char* want_name, SP& kernel_body,
                std::string& sig_text) {
    size_t i = 0;
    while (i < T.size()) {
      const Tok& t = T[i];
      if (t.s == "typedef") {
        size_t j = i + 1;
        std::string ty;
        while (j < T.size() && T[j].s != ";") ty += (ty.empty() ? "" : " ") + T[j++].s;
        if (ty.size() > 6 && ty.compare(ty.size() - 6, 6, " int_t") == 0) {
          std::string base = ty.substr(0, ty.size() - 6);
          if (base == "long" || base == "long long" || base == "int64_t") k.int_size = 8;
          else if (base == "int" || base == "int32_t") k.int_size = 4;
          else fail(DGK_ERR_UNSUPPORTED, t.line, "int_t as %s", base.c_str());
          have_int_t = true;
        }
        i = j + 1;
        continue;
      }
      if (t.s == "const" && i + 5 < T.size() && Parser::is_type(T[i + 1].s) && T[i + 2].k == T_ID &&
          T[i + 3].s == "[") {
        read_static(T, i);
        continue;
      }
      if (t.k == T_ATTR && t.s == "@kernel") {
        if (i + 3 >= T.size() || T[i + 1].s != "void" || T[i + 2].k != T_ID || T[i + 3].s != "(")
          fail(DGK_ERR_INVALID, t.line, "@kernel void name(...) expected");
        std::string nm = T[i + 2].s;
        size_t j = i + 4;
        int dpt = 0;
        std::string sig;
        for (; j < T.size(); ++j) {
          if (T[j].s == "(") ++dpt;
          if (T[j].s == ")") {
            if (dpt == 0) break;
            --dpt;
          }
          sig += T[j].s + " ";
        }
        if (j >= T.size()) fail(DGK_ERR_INVALID, t.line, "unterminated kernel signature");
        StmtParser sp(T, j + 1);
        SP body = sp.block();
        if (nm == want_name) {
          if (kernel_body) fail(DGK_ERR_INVALID, t.line, "kernel %s defined twice", want_name);
          kernel_body = body;
          sig_text = sig;
        }
        i = sp.i;
        continue;
      }
      // anything else at top level: skip one declaration or block
      int dpt = 0;
      for (; i < T.size(); ++i) {
        if (T[i].s == "{") ++dpt;
        if (T[i].s == "}") {
          if (--dpt == 0) {
            ++i;
            break;
          }
        }
        if (T[i].s == ";" && dpt == 0) {
          ++i;
          break;
        }
      }
    }
  }

  void read_static(const std::vector<Tok>& T, size_t& i) {
    const int line = T[i].line;
    std::string ty = T[i + 1].s, name = T[i + 2].s;
    size_t j = i + 3;
    long long dims[2];
    for (int d = 0; d < 2; ++d) {
      if (j + 2 >= T.size() || T[j].s != "[" || T[j + 1].k != T_NUM || T[j + 2].s != "]")
        fail(DGK_ERR_UNSUPPORTED, line, "static array %s: two constant dimensions expected", name.c_str());
      dims[d] = atoll(T[j + 1].s.c_str());
      j += 3;
    }
    if (j >= T.size() || T[j].s != "=") fail(DGK_ERR_UNSUPPORTED, line, "static array %s without values", name.c_str());
    ++j;
    int et = elem_type_of(ty, k.int_size, line);
    if (et != ET_F64 && et != ET_F32) fail(DGK_ERR_UNSUPPORTED, line, "static array of %s", ty.c_str());
    std::vector<double> vals;
    int dpt = 0;
    bool neg = false;
    for (; j < T.size(); ++j) {
      const std::string& s = T[j].s;
      if (s == "{") ++dpt;
      else if (s == "}") {
        if (--dpt == 0) break;
      } else if (s == "-") neg = !neg;
      else if (s == "+" || s == ",") {
      } else if (T[j].k == T_NUM) {
        std::string v = s;
        while (!v.empty() && (v.back() == 'f' || v.back() == 'F')) v.pop_back();
        double x = et == ET_F32 ? (double)strtof(v.c_str(), nullptr) : strtod(v.c_str(), nullptr);
        vals.push_back(neg ? -x : x);
        neg = false;
      } else {
        fail(DGK_ERR_UNSUPPORTED, T[j].line, "non-literal value in static array %s", name.c_str());
      }
    }
    if (j + 1 >= T.size() || T[j + 1].s != ";") fail(DGK_ERR_INVALID, line, "static array %s not terminated", name.c_str());
    if ((long long)vals.size() != dims[0] * dims[1])
      fail(DGK_ERR_INVALID, line, "static array %s holds %d values for [%lld][%lld]", name.c_str(),
           (int)vals.size(), dims[0], dims[1]);
    static_vals[name] = vals;  // [cols][rows]: column-major
    static_dims[name] = {(int)dims[0], (int)dims[1]};
    static_type[name] = et;
    i = j + 2;
  }

  void read_signature(const std::string& sig, int line) {
    std::vector<Tok> t;
    tokenize(sig, line, t);
    std::vector<std::vector<Tok>> parts(1);
    for (auto& x : t) {
      if (x.s == ",") parts.emplace_back();
      else parts.back().push_back(x);
    }
    for (auto& p : parts) {
      if (p.empty()) continue;
      Param a;
      const size_t last = p.size() - 1;
      if (p[last].k != T_ID || Parser::is_type(p[last].s))
        fail(DGK_ERR_INVALID, line, "cannot read kernel parameter '%s'", p[0].s.c_str());
      a.name = p[last].s;
      for (size_t q = 0; q < last; ++q) {
        if (p[q].s == "const") a.is_const = true;
        else if (p[q].s == "*") a.ptr = true;
        else if (p[q].k == T_ID && Parser::is_type(p[q].s)) a.type += (a.type.empty() ? "" : " ") + p[q].s;
        else if (p[q].s != "__restrict__" && p[q].s != "restrict")
          fail(DGK_ERR_INVALID, line, "cannot read kernel parameter '%s'", a.name.c_str());
      }
      if (a.type.empty()) fail(DGK_ERR_INVALID, line, "parameter %s has no type", a.name.c_str());
      k.params.push_back(a);
    }
    if (k.params.empty() || k.params[0].name != "K" || !k.params[0].ptr)
      fail(DGK_ERR_INVALID, line, "the kernel's first parameter must be the K array "
           "(runner/kernel_execution.go:150-158)");
    if (elem_type_of(k.params[0].type, k.int_size, line) != (k.int_size == 8 ? ET_I64 : ET_I32))
      fail(DGK_ERR_INVALID, line, "K must be int_t*");
    for (size_t i = 0; i < k.params.size(); ++i) {
      const Param& a = k.params[i];
      const std::string& n = a.name;
      if (a.ptr && n.size() > 7 && n.compare(n.size() - 7, 7, "_global") == 0) {
        Array arr;
        arr.name = n.substr(0, n.size() - 7);
        arr.gparam = (int)i;
        arr.type = elem_type_of(a.type, k.int_size, line);
        arr.is_const = a.is_const;
        int o = param_pos(arr.name + "_offsets");
        if (o < 0 || !k.params[o].ptr) fail(DGK_ERR_INVALID, line, "%s has no %s_offsets", n.c_str(), arr.name.c_str());
        if (elem_type_of(k.params[o].type, k.int_size, line) != (k.int_size == 8 ? ET_I64 : ET_I32))
          fail(DGK_ERR_INVALID, line, "%s_offsets must be int_t*", arr.name.c_str());
        arr.oparam = o;
        auto pt = part_type.find(arr.name);
        if (pt != part_type.end() && pt->second != arr.type)
          fail(DGK_ERR_INVALID, line, "%s_PART casts to %s but %s is %s", arr.name.c_str(),
               et_name(pt->second), n.c_str(), a.type.c_str());
        k.arrays.push_back(arr);
      }
    }
  }

  void build(const char* src, const char* name) {
    Prep p = preprocess(src);
    macros = p.macros;
    auto num_macro = [&](const char* n, int& v) {
      auto it = macros.find(n);
      if (it == macros.end() || it->second.fn || it->second.body.size() != 1 ||
          it->second.body[0].k != T_NUM)
        fail(DGK_ERR_INVALID, 0, "the preamble's #define %s is missing (runner/builder/builder.go:193-211)", n);
      v = atoi(it->second.body[0].s.c_str());
    };
    num_macro("NPART", k.npart);
    num_macro("KpartMax", k.kpart_max);
    if (k.npart < 1) fail(DGK_ERR_INVALID, 0, "NPART %d", k.npart);
    std::vector<Tok> toks;
    std::set<std::string> active;
    // int_t first: the statement reader needs the width
    for (size_t i = 0; i + 2 < p.toks.size(); ++i)
      if (p.toks[i].s == "typedef" && p.toks[i + 2].s == "int_t")
        k.int_size = (p.toks[i + 1].s == "int") ? 4 : 8;
    expand(p.toks, macros, active, toks, 0);
    read_macro_info();
    SP body;
    std::string sig;
    read_top(toks, name, body, sig);
    if (!have_int_t) fail(DGK_ERR_INVALID, 0, "the preamble's typedef of int_t is missing");
    if (!body) fail(DGK_ERR_NOT_FOUND, 0, "no @kernel named %s in the source", name);
    read_signature(sig, body->line);
    std::vector<SP> top = stmts_of(body);
    if (top.size() != 1 || top[0]->k != Stmt::FOR || top[0]->attr != "@outer")
      fail(DGK_ERR_UNSUPPORTED, body->line, "kernel body must be one @outer loop over partitions");
    const SP& outer = top[0];
    EP ob;
    loop_shape(*outer, ob);
    long long np;
    if (!(is_id(ob, "NPART") || (const_eval(ob, np) && np == k.npart)))
      fail(DGK_ERR_UNSUPPORTED, outer->line, "@outer loop must run over NPART partitions");
    part_var = outer->var;
    read_part_body(stmts_of(outer->body[0]));
    if (k.ops.empty()) fail(DGK_ERR_UNSUPPORTED, body->line, "kernel does nothing this library runs");
    fuse();
  }
};

std::string describe(const dgk_kernel_s& k) {
  std::string s = "kernel " + k.name + ": NPART " + std::to_string(k.npart) + ", KpartMax " +
                  std::to_string(k.kpart_max) + ", int_t " + (k.int_size == 8 ? "long" : "int") + "\n";
  for (const Op& o : k.ops) {
    switch (o.k) {
      case OP_MATMUL: {
        const Matrix& M = k.mats[o.mat];
        s += std::string(o.add ? "matmul_add " : "matmul ") + M.name + " [" + std::to_string(M.rows) +
             "x" + std::to_string(M.cols) + " " + (M.is_static ? "static" : "device") + " " +
             et_name(M.type) + (o.rs == 1 ? ", column-major" : ", row-major") + ", sum " +
             et_name(o.atype) + "] " + k.arrays[o.in].name + " -> " + k.arrays[o.out].name + "\n";
        break;
      }
      case OP_PW:
        s += "pointwise " + o.text + "\n";
        break;
      case OP_DERIV:
        s += o.text + "\n";
        break;
    }
  }
  return s;
}

int build_kernel(dgk_kernel_s& k, const char* source, const char* name) {
  try {
    Builder b(k);
    b.build(source, name);
  } catch (const Fail& f) {
    dgk::set_error("%s: %s", name, f.msg.c_str());
    return f.code;
  } catch (const std::exception& e) {
    dgk::set_error("%s: %s", name, e.what());
    return DGK_ERR_INVALID;
  }
  return DGK_OK;
}

// ---------------------------------------------------------------------------
// Launch
// ---------------------------------------------------------------------------
struct Bound {
  const dgk_arg* args;
  const dgk_kernel_s* k;
  const void* G(int a) const { return args[k->arrays[a].gparam].v.mem; }
  const void* O(int a) const { return args[k->arrays[a].oparam].v.mem; }
};

int run_pw(dgk_kernel_s& k, const Op& o, const Bound& b, const int64_t* K, const void* kdev) {
  PwProg pg{};
  pg.nops = (int)o.code.size();
  pg.per_elem = o.per_elem;
  pg.off_size = k.int_size;
  for (size_t i = 0; i < o.code.size(); ++i) pg.op[i] = o.code[i];
  for (size_t i = 0; i < o.imm.size(); ++i) {
    int p = o.imm_param[i];
    if (p < 0) {
      pg.imm[i] = o.imm[i];
      continue;
    }
    const dgk_arg& a = b.args[p];
    switch (a.kind) {
      case DGK_ARG_F64: pg.imm[i] = a.v.f64; break;
      case DGK_ARG_F32: pg.imm[i] = (double)a.v.f32; break;
      case DGK_ARG_I64: pg.imm[i] = (double)a.v.i64; break;
      case DGK_ARG_I32: pg.imm[i] = (double)a.v.i32; break;
      default: break;
    }
  }
  for (size_t i = 0; i < o.arr.size(); ++i) {
    pg.base[i] = b.G(o.arr[i]);
    pg.off[i] = b.O(o.arr[i]);
    pg.type[i] = k.arrays[o.arr[i]].type;
  }
  int per_tile = std::max(1, 1024 / o.per_elem);
  dgk::PartTiles pt;
  dim3 grid;
  int rc = dgk::build_part_tiles(k.dev, k.npart, K, per_tile, kdev, k.int_size, pt, grid);
  if (rc) return rc;
  if (grid.x == 0) return DGK_OK;
  hipLaunchKernelGGL(pointwise_kernel, grid, dim3(256), 0, k.dev->stream, pt, per_tile, pg);
  DGK_LAUNCH_CHECK();
  return DGK_OK;
}

const void* matrix_ptr(const dgk_kernel_s& k, int m, const dgk_arg* args) {
  const Matrix& M = k.mats[m];
  return M.is_static ? M.dev : args[M.param].v.mem;
}

int dgk_type(int et) { return et == ET_F32 ? DGK_F32 : DGK_F64; }

}  // namespace

extern "C" int dgk_kernel_list(char* buf, size_t size) {
  const char* forms =
      "alias: [const] T* A = X_PART(part);\n"
      "matmul: MATMUL_X(IN, OUT, K[part]); MATMUL_ADD_X(IN, OUT, K[part]);\n"
      "element-node loop: for (k < KpartMax; @inner) if (k < K[part]) for (n < C) { int i = n + k*C; A[i] = expr; }\n"
      "flat-node loop: for (i < C*KpartMax; @inner) { A[i] = expr; }\n"
      "element matvec: for (e < K[part]) for (i < R; @inner) { T s = 0; for (j < C) s += M[i*C + j] * U[e*C + j]; V[e*R + i] = s; }\n"
      "fused: MATMUL of three Np x Np operators + chain rule with nine metric arrays (physical derivative)\n";
  DGK_REQUIRE(buf && size > strlen(forms), "dgk_kernel_list: buffer too small");
  memcpy(buf, forms, strlen(forms) + 1);
  return DGK_OK;
}

extern "C" int dgk_kernel_describe(const char* source, const char* name, char* buf, size_t size) {
  dgk::clear_error();
  DGK_REQUIRE(source && name && buf && size > 0, "dgk_kernel_describe: NULL argument");
  dgk_kernel_s k;
  k.name = name;
  int rc = build_kernel(k, source, name);
  if (rc) return rc;
  std::string d = describe(k);
  DGK_REQUIRE(size > d.size(), "dgk_kernel_describe: buffer of %zu bytes, need %zu", size, d.size() + 1);
  memcpy(buf, d.c_str(), d.size() + 1);
  return DGK_OK;
}

extern "C" int dgk_kernel_build_from_source(dgk_device dev, const char* source, const char* name,
                                            dgk_kernel* out) {
  dgk::clear_error();
  DGK_REQUIRE(dev && source && name && out, "dgk_kernel_build_from_source: NULL argument");
  *out = nullptr;
  auto* k = new dgk_kernel_s();
  k->dev = dev;
  k->name = name;
  int rc = build_kernel(*k, source, name);
  if (rc == DGK_OK) {
    for (Matrix& M : k->mats) {
      if (!M.is_static) continue;
      size_t n = M.colmajor.size();
      if (M.type == ET_F32) {
        std::vector<float> f(n);
        for (size_t i = 0; i < n; ++i) f[i] = (float)M.colmajor[i];
        rc = dgk_malloc(dev, n * 4, f.data(), &M.dev);
      } else {
        rc = dgk_malloc(dev, n * 8, M.colmajor.data(), &M.dev);
      }
      if (rc) break;
    }
  }
  if (rc) {
    std::string msg = dgk_last_error();
    dgk_kernel_free(k);
    dgk::set_error("%s", msg.c_str());
    return rc;
  }
  *out = k;
  return DGK_OK;
}

extern "C" int dgk_kernel_launch(dgk_kernel k, const void* host_K, int nargs, const dgk_arg* args) {
  dgk::clear_error();
  DGK_REQUIRE(k && args, "dgk_kernel_launch: NULL argument");
  DGK_REQUIRE(nargs == (int)k->params.size(), "%s: %d arguments, the signature has %d", k->name.c_str(),
              nargs, (int)k->params.size());
  // every argument against its parameter
  for (int i = 0; i < nargs; ++i) {
    const Param& p = k->params[i];
    const dgk_arg& a = args[i];
    if (p.ptr) {
      DGK_REQUIRE(a.kind == DGK_ARG_MEM && a.v.mem, "%s: argument %d (%s) must be non-NULL device memory",
                  k->name.c_str(), i, p.name.c_str());
      continue;
    }
    int et = ET_F64;
    try {
      et = elem_type_of(p.type, k->int_size, 0);
    } catch (const Fail& f) {
      dgk::set_error("%s: %s", k->name.c_str(), f.msg.c_str());
      return f.code;
    }
    int want = et == ET_F64 ? DGK_ARG_F64 : et == ET_F32 ? DGK_ARG_F32 : et == ET_I32 ? DGK_ARG_I32
                                                                                    : DGK_ARG_I64;
    DGK_REQUIRE(a.kind == want, "%s: scalar %s (%s) passed with kind %d, expected %d", k->name.c_str(),
                p.name.c_str(), p.type.c_str(), a.kind, want);
  }
  const void* kdev = args[0].v.mem;
  std::vector<int64_t> K(k->npart);
  if (host_K) {
    for (int p = 0; p < k->npart; ++p)
      K[p] = k->int_size == 8 ? ((const int64_t*)host_K)[p] : ((const int32_t*)host_K)[p];
  } else {
    std::vector<char> raw((size_t)k->npart * k->int_size);
    DGK_HIP(hipMemcpyAsync(raw.data(), kdev, raw.size(), hipMemcpyDeviceToHost, k->dev->stream));
    DGK_HIP(hipStreamSynchronize(k->dev->stream));
    for (int p = 0; p < k->npart; ++p)
      K[p] = k->int_size == 8 ? ((const int64_t*)raw.data())[p] : ((const int32_t*)raw.data())[p];
  }
  for (int p = 0; p < k->npart; ++p)
    DGK_REQUIRE(K[p] >= 0 && K[p] <= k->kpart_max, "%s: K[%d] = %lld outside 0..KpartMax (%d)",
                k->name.c_str(), p, (long long)K[p], k->kpart_max);
  Bound b{args, k};
  for (const Op& o : k->ops) {
    int rc = DGK_OK;
    switch (o.k) {
      case OP_MATMUL: {
        const Matrix& M = k->mats[o.mat];
        rc = dgk::matmul_typed(k->dev, k->npart, K.data(), kdev, k->int_size, M.rows, M.cols,
                               matrix_ptr(*k, o.mat, args), o.rs, o.cs, dgk_type(M.type), b.G(o.in),
                               b.O(o.in), dgk_type(k->arrays[o.in].type), (void*)b.G(o.out), b.O(o.out),
                               dgk_type(k->arrays[o.out].type), dgk_type(o.atype), o.add, k->int_size);
        break;
      }
      case OP_PW:
        rc = run_pw(*k, o, b, K.data(), kdev);
        break;
      case OP_DERIV: {
        const double* D[3];
        const void* mets[9];
        const void* moff[9];
        void* outs[3];
        const void* ooff[3];
        double* tmp[3];
        const void* toff[3];
        for (int q = 0; q < 3; ++q) {
          D[q] = (const double*)matrix_ptr(*k, o.dmat[q], args);
          outs[q] = (void*)b.G(o.outs[q]);
          ooff[q] = b.O(o.outs[q]);
          tmp[q] = (double*)b.G(o.temps[q]);
          toff[q] = b.O(o.temps[q]);
        }
        for (int q = 0; q < 9; ++q) {
          mets[q] = b.G(o.mets[q]);
          moff[q] = b.O(o.mets[q]);
        }
        const int np = k->mats[o.dmat[0]].rows;
        rc = dgk::phys_deriv_nodal_full(k->dev, np_order(np), k->npart, K.data(), kdev, k->int_size, D, 1,
                                        np, (const double*)b.G(o.u), b.O(o.u), mets, moff,
                                        dgk_type(k->arrays[o.mets[0]].type), outs, ooff,
                                        dgk_type(k->arrays[o.outs[0]].type), tmp, toff, k->int_size,
                                        false);
        break;
      }
    }
    if (rc) return rc;
  }
  return DGK_OK;
}

extern "C" int dgk_kernel_free(dgk_kernel k) {
  if (!k) return DGK_OK;
  for (Matrix& M : k->mats)
    if (M.dev) dgk_free(k->dev, M.dev);
  delete k;
  return DGK_OK;
}
