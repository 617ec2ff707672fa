// Constraint evaluation on the quotient domain: AIR programs compiled to gfx950 kernels.
//
// Replaces evaluate_constraints_into (p3-miden-lifted-stark/src/prover/constraints/
// mod.rs:63-179): at every point x_i = shift w^i of the quotient domain gJ (natural
// order, |gJ| = n d), evaluate the AIR's constraints on the current and next trace
// rows, fold them as sum_k alpha^(K-1-k) C_k, and add the result to the accumulator.
//
// The AIR arrives as a straight-line program (include/p3gpu.h, p3g_air_op), the form
// a symbolic AIR builder records.  This file checks it, infers which values are base
// and which extension field elements, and writes it out as the body of a kernel:
// one typed local per value, base-field arithmetic where both operands are base, and
// the fold as multiply-adds of each constraint with a precomputed power of alpha into
// two unreduced 192-bit sums (one reduction per point).  hiprtc compiles the kernel
// once per program (air_kernel.hpp holds its fixed part: point tiling, selectors,
// the accumulator update); the code object is cached for the process and loaded once
// per device.
#include <hip/hiprtc.h>

#include <map>
#include <memory>
#include <mutex>
#include <sstream>
#include <string>

#include "internal.hpp"
#include "air_kernel.hpp"

namespace p3g {

namespace {

// the headers the generated source includes, as text (generated at build time)
const char kGlHpp[] =
#include "gl_hpp.inc"
    ;
const char kAirKernelHpp[] =
#include "air_kernel_hpp.inc"
    ;

enum AirOp : uint32_t {
  OP_CONST = 0,
  OP_MAIN = 1,
  OP_AUX = 2,
  OP_PERIODIC = 3,
  OP_PUBLIC = 4,
  OP_EXT_INPUT = 5,
  OP_FIRST = 6,
  OP_LAST = 7,
  OP_TRANSITION = 8,
  OP_ADD = 9,
  OP_SUB = 10,
  OP_MUL = 11,
  OP_NEG = 12,
  OP_ASSERT = 13,
  OP_ASSERT_EXT = 14,
};

constexpr uint64_t AIR_MAX_OPS = uint64_t(1) << 20;
constexpr unsigned AIR_MAX_LOG_D = 20;
constexpr uint64_t AIR_BLOCK = 128;   // ops per basic block of a compiled program
constexpr uint64_t AIR_FUNC = 1024;   // ops per device function (a multiple of AIR_BLOCK)

bool is_assert(uint32_t op) { return op == OP_ASSERT || op == OP_ASSERT_EXT; }

// What a program reads and what each of its values is.
struct AirShape {
  std::vector<uint8_t> ext;   // 1: the value is an extension element
  std::vector<uint8_t> used;  // 1: a later op reads the value
  uint64_t nasserts = 0;
  bool selectors = false, aux = false, periodic = false;
  uint64_t main_cols = 0, aux_cols = 0, per_cols = 0, npub = 0, next = 0;  // 1 + the largest index read
};

// Checks that do not depend on the matrices: op codes, operands naming earlier values
// (never an assertion), canonical constants, at least one assertion.
AirShape analyze(const AirOpIn* ops, uint64_t nops) {
  if (nops == 0) throw ApiError(E_EMPTY, "empty constraint program");
  if (nops > AIR_MAX_OPS) throw ApiError(E_RANGE, "constraint program longer than 2^20 ops");
  AirShape s;
  s.ext.assign(nops, 0);
  s.used.assign(nops, 0);
  auto operand = [&](uint64_t k, uint32_t v) {
    if (v >= k || is_assert(ops[v].op))
      throw ApiError(E_INVAL, "constraint program: op " + std::to_string(k) + " reads an undefined value");
    s.used[v] = 1;
    return s.ext[v];
  };
  auto upto = [](uint64_t& m, uint32_t v) { m = std::max<uint64_t>(m, uint64_t(v) + 1); };
  for (uint64_t k = 0; k < nops; k++) {
    const auto& o = ops[k];
    switch (o.op) {
      case OP_CONST:
        if (o.imm >= P) throw ApiError(E_INVAL, "constraint program: non-canonical constant");
        break;
      case OP_MAIN:
        if (o.b > 1) throw ApiError(E_INVAL, "constraint program: row must be 0 (current) or 1 (next)");
        upto(s.main_cols, o.a);
        break;
      case OP_AUX:
        if (o.b > 1) throw ApiError(E_INVAL, "constraint program: row must be 0 (current) or 1 (next)");
        upto(s.aux_cols, o.a);
        s.aux = true;
        s.ext[k] = 1;
        break;
      case OP_PERIODIC:
        upto(s.per_cols, o.a);
        s.periodic = true;
        break;
      case OP_PUBLIC: upto(s.npub, o.a); break;
      case OP_EXT_INPUT:
        upto(s.next, o.a);
        s.ext[k] = 1;
        break;
      case OP_FIRST:
      case OP_LAST:
      case OP_TRANSITION: s.selectors = true; break;
      case OP_ADD:
      case OP_SUB:
      case OP_MUL: {
        const uint8_t ea = operand(k, o.a), eb = operand(k, o.b);
        s.ext[k] = ea | eb;
        break;
      }
      case OP_NEG: s.ext[k] = operand(k, o.a); break;
      case OP_ASSERT:
      case OP_ASSERT_EXT:
        operand(k, o.a);
        s.nasserts++;
        break;
      default: throw ApiError(E_INVAL, "constraint program: unknown op " + std::to_string(o.op));
    }
  }
  if (s.nasserts == 0) throw ApiError(E_EMPTY, "constraint program asserts nothing");
  return s;
}

bool is_leaf(uint32_t op) { return op <= OP_TRANSITION; }

// The value of one op, in the names of a chunk function's parameters (mc, mn: main rows;
// ac, an: aux rows; per: periodic row; pub, ext: inputs; first, last, trans: selectors)
// and of earlier values (v<index>).
std::string air_expr(const AirOpIn& op, const AirShape& s) {
  std::ostringstream o;
  const std::string x = "v" + std::to_string(op.a), y = "v" + std::to_string(op.b);
  const bool ex = op.op >= OP_ADD && s.ext[op.a], ey = op.op >= OP_ADD && op.op <= OP_MUL && s.ext[op.b];
  switch (op.op) {
    case OP_CONST: o << "0x" << std::hex << op.imm << std::dec << "ull"; break;
    case OP_MAIN: o << (op.b ? "mn[" : "mc[") << op.a << "]"; break;
    case OP_AUX: {
      const char* row = op.b ? "an" : "ac";
      o << "ef2{" << row << "[" << 2ull * op.a << "], " << row << "[" << 2ull * op.a + 1 << "]}";
      break;
    }
    case OP_PERIODIC: o << "per[" << op.a << "]"; break;
    case OP_PUBLIC: o << "pub[" << op.a << "]"; break;
    case OP_EXT_INPUT: o << "ext[" << op.a << "]"; break;
    case OP_FIRST: o << "first"; break;
    case OP_LAST: o << "last"; break;
    case OP_TRANSITION: o << "trans"; break;
    case OP_ADD:
      if (ex && ey) o << "ef_add(" << x << ", " << y << ")";
      else if (ex) o << "ef2{gl_add(" << x << ".c0, " << y << "), " << x << ".c1}";
      else if (ey) o << "ef2{gl_add(" << x << ", " << y << ".c0), " << y << ".c1}";
      else o << "gl_add(" << x << ", " << y << ")";
      break;
    case OP_SUB:
      if (ex && ey) o << "ef_sub(" << x << ", " << y << ")";
      else if (ex) o << "ef2{gl_sub(" << x << ".c0, " << y << "), " << x << ".c1}";
      else if (ey) o << "ef2{gl_sub(" << x << ", " << y << ".c0), gl_neg(" << y << ".c1)}";
      else o << "gl_sub(" << x << ", " << y << ")";
      break;
    case OP_MUL:
      if (ex && ey) o << "ef_mul(" << x << ", " << y << ")";
      else if (ex) o << "ef_mul_base(" << x << ", " << y << ")";
      else if (ey) o << "ef_mul_base(" << y << ", " << x << ")";
      else o << "gl_mul(" << x << ", " << y << ")";
      break;
    case OP_NEG: o << (ex ? "ef_neg(" : "gl_neg(") << x << ")"; break;
  }
  return o.str();
}

// The kernel source of a program, in two levels:
//   - blocks of AIR_BLOCK consecutive ops, each a basic block of its own (entered on a
//     kernel-argument bit that is always set): the compiler's instruction selection and
//     scheduling work on one block at a time.  Values read by a later block are declared
//     at function scope and stay in registers; a block re-emits the leaves (loads, inputs,
//     constants, selectors) it reads from elsewhere.
//   - functions of AIR_FUNC ops (whole blocks), which bound the work of register
//     allocation.  A computed value read by a later function goes through a per-thread
//     slot array (slots reused once their last reader has run).  Each function returns
//     its constraints folded with their powers of alpha; air_point sums them.  A program
//     of one function is inlined whole.
// Compile time is linear in the program; the calls and slots cost little next to the
// arithmetic of AIR_FUNC ops.
std::string air_source(const AirOpIn* ops, uint64_t nops, const AirShape& s) {
  const uint64_t nfn = (nops + AIR_FUNC - 1) / AIR_FUNC;
  auto fn_of = [](uint64_t k) { return k / AIR_FUNC; };
  auto blk_of = [](uint64_t k) { return k / AIR_BLOCK; };
  auto operands = [&](uint64_t k, auto&& f) {
    const uint32_t op = ops[k].op;
    if (op == OP_ADD || op == OP_SUB || op == OP_MUL) {
      f(ops[k].a);
      f(ops[k].b);
    } else if (op == OP_NEG || is_assert(op)) {
      f(ops[k].a);
    }
  };
  const uint64_t nblk = (nops + AIR_BLOCK - 1) / AIR_BLOCK;
  // readers: the last function and block reading each value; what each function reads
  // from earlier functions (computed values) and each block from elsewhere (leaves)
  std::vector<uint64_t> last_fn(nops, 0), last_blk(nops, 0), seen_fn(nops, ~uint64_t(0)),
      seen_blk(nops, ~uint64_t(0));
  std::vector<std::vector<uint32_t>> fn_imports(nfn), blk_leaves(nblk);
  for (uint64_t k = 0; k < nops; k++) {
    const uint64_t f = fn_of(k), b = blk_of(k);
    operands(k, [&](uint32_t v) {
      last_fn[v] = std::max(last_fn[v], f);
      last_blk[v] = std::max(last_blk[v], b);
      if (is_leaf(ops[v].op)) {
        if (blk_of(v) != b && seen_blk[v] != b) {
          seen_blk[v] = b;
          blk_leaves[b].push_back(v);
        }
      } else if (fn_of(v) < f && seen_fn[v] != f) {
        seen_fn[v] = f;
        fn_imports[f].push_back(v);
      }
    });
  }
  // slots (two words each) for computed values read in a later function
  std::vector<int64_t> slot(nops, -1);
  std::vector<std::vector<uint32_t>> frees(nfn), exports(nfn);
  std::vector<uint32_t> free_slots;
  uint32_t nslots = 0;
  for (uint64_t f = 0; f < nfn; f++) {
    for (uint32_t v : frees[f]) free_slots.push_back((uint32_t)slot[v]);  // read at the top of f
    for (uint64_t k = f * AIR_FUNC; k < std::min(nops, (f + 1) * AIR_FUNC); k++) {
      if (is_leaf(ops[k].op) || is_assert(ops[k].op) || last_fn[k] <= f) continue;
      if (free_slots.empty()) free_slots.push_back(nslots++);
      slot[k] = free_slots.back();
      free_slots.pop_back();
      frees[last_fn[k]].push_back((uint32_t)k);
      exports[f].push_back((uint32_t)k);
    }
  }

  std::ostringstream o;
  o << "#define AIR_SELECTORS " << int(s.selectors) << "\n"
    << "#define AIR_AUX " << int(s.aux) << "\n"
    << "#define AIR_PERIODIC " << int(s.periodic) << "\n"
    << "#include \"gl.hpp\"\n"
    << "#include \"air_kernel.hpp\"\n"
    << "namespace p3g {\n";
  auto type = [&](uint64_t k) { return s.ext[k] ? "ef2 " : "uint64_t "; };
  uint64_t cons = 0;  // constraints emitted
  for (uint64_t f = 0; f < nfn; f++) {
    const uint64_t k0 = f * AIR_FUNC, k1 = std::min(nops, (f + 1) * AIR_FUNC);
    o << "__device__ " << (nfn > 1 ? "__noinline__" : "__forceinline__") << " ef2 air_fn_" << f
      << "(const uint64_t* __restrict__ ap, const uint64_t* __restrict__ pub, const ef2* __restrict__ ext,"
      << " const uint64_t* mc, const uint64_t* mn, const uint64_t* ac, const uint64_t* an, const uint64_t* per,"
      << " uint64_t first, uint64_t last, uint64_t trans, uint64_t* live, uint32_t on) {\n"
      << "  Acc192 s0 = acc192_zero(), s1 = acc192_zero();\n";
    for (uint32_t v : fn_imports[f]) {
      if (s.ext[v]) o << "  const ef2 v" << v << " = ef2{live[" << 2 * slot[v] << "], live[" << 2 * slot[v] + 1 << "]};\n";
      else o << "  const uint64_t v" << v << " = live[" << 2 * slot[v] << "];\n";
    }
    // computed values of this function read outside their own block
    std::vector<uint8_t> fscope(k1 - k0, 0);
    for (uint64_t k = k0; k < k1; k++) {
      if (is_leaf(ops[k].op) || is_assert(ops[k].op) || !s.used[k] || last_blk[k] <= blk_of(k)) continue;
      fscope[k - k0] = 1;
      o << "  " << type(k) << "v" << k << ";\n";
    }
    const bool guard = k1 - k0 > AIR_BLOCK;  // a lone block needs no guard
    for (uint64_t b = blk_of(k0); b * AIR_BLOCK < k1; b++) {
      o << (guard ? "  if (on >> " + std::to_string(b % 32) + " & 1) {\n" : "  {\n");
      for (uint32_t v : blk_leaves[b]) o << "  const " << type(v) << "v" << v << " = " << air_expr(ops[v], s) << ";\n";
      for (uint64_t k = b * AIR_BLOCK; k < std::min(k1, (b + 1) * AIR_BLOCK); k++) {
        const auto& op = ops[k];
        if (is_assert(op.op)) {
          const uint64_t pw = 3 * (s.nasserts - 1 - cons++);  // alpha^(K-1-c) as (c0, c1, 7 c1)
          const std::string x = "v" + std::to_string(op.a);
          if (s.ext[op.a]) {
            // (x0 + x1 X)(a0 + a1 X) = (x0 a0 + x1 (7 a1)) + (x1 a0 + x0 a1) X
            o << "  acc192_mad(s0, " << x << ".c0, ap[" << pw << "]); acc192_mad(s0, " << x << ".c1, ap["
              << pw + 2 << "]); acc192_mad(s1, " << x << ".c1, ap[" << pw << "]); acc192_mad(s1, " << x
              << ".c0, ap[" << pw + 1 << "]);\n";
          } else {
            o << "  acc192_mad(s0, " << x << ", ap[" << pw << "]); acc192_mad(s1, " << x << ", ap[" << pw + 1
              << "]);\n";
          }
        } else if (s.used[k]) {
          if (fscope[k - k0]) o << "  v" << k << " = " << air_expr(op, s) << ";\n";
          else o << "  const " << type(k) << "v" << k << " = " << air_expr(op, s) << ";\n";
        }
      }
      o << "  }\n";
    }
    for (uint32_t v : exports[f]) {
      if (s.ext[v]) o << "  live[" << 2 * slot[v] << "] = v" << v << ".c0; live[" << 2 * slot[v] + 1 << "] = v" << v << ".c1;\n";
      else o << "  live[" << 2 * slot[v] << "] = v" << v << ";\n";
    }
    o << "  return ef2{acc192_reduce(s0), acc192_reduce(s1)};\n}\n";
  }
  o << "__device__ __forceinline__ ef2 air_point(const AirKArgs& a, const AirRow& r) {\n";
  if (nslots) o << "  uint64_t live[" << 2 * nslots << "];\n";
  else o << "  uint64_t* live = nullptr;\n";
  for (uint64_t f = 0; f < nfn; f++) {
    const std::string call = "air_fn_" + std::to_string(f) +
                             "(a.apow, a.pub, a.ext, r.mc, r.mn, r.ac, r.an, r.per, r.first, r.last, r.trans, live, "
                             "a.blocks_on)";
    o << (f ? "  f = ef_add(f, " + call + ");\n" : "  ef2 f = " + call + ";\n");
  }
  o << "  return f;\n}\n}  // namespace p3g\n";
  return o.str();
}

std::mutex g_air_mu;
std::map<std::string, std::shared_ptr<const std::string>> g_air_code;        // program -> code object
std::map<std::pair<int, const std::string*>, hipFunction_t> g_air_function;  // (device, code) -> kernel

void rtc_check(hiprtcResult r, const char* what) {
  if (r != HIPRTC_SUCCESS) throw ApiError(E_DEVICE, std::string(what) + ": " + hiprtcGetErrorString(r));
}

// The compiled kernel of a program (compiled on first use, then cached).
std::shared_ptr<const std::string> air_code(const AirOpIn* ops, uint64_t nops, const AirShape& s) {
  std::string key(reinterpret_cast<const char*>(ops), nops * sizeof(AirOpIn));
  {
    std::lock_guard<std::mutex> g(g_air_mu);
    auto it = g_air_code.find(key);
    if (it != g_air_code.end()) return it->second;
  }
  const std::string src = air_source(ops, nops, s);
  hiprtcProgram prog;
  const char* hdr_src[] = {kGlHpp, kAirKernelHpp};
  const char* hdr_name[] = {"gl.hpp", "air_kernel.hpp"};
  rtc_check(hiprtcCreateProgram(&prog, src.c_str(), "air.hip", 2, hdr_src, hdr_name), "hiprtcCreateProgram");
  const char* opts[] = {"--offload-arch=gfx950", "-O3", "-std=c++17",
                        "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"};
  const hiprtcResult r = hiprtcCompileProgram(prog, (int)(sizeof(opts) / sizeof(opts[0])), opts);
  if (r != HIPRTC_SUCCESS) {
    size_t n = 0;
    std::string log;
    if (hiprtcGetProgramLogSize(prog, &n) == HIPRTC_SUCCESS && n) {
      log.resize(n);
      hiprtcGetProgramLog(prog, &log[0]);
    }
    hiprtcDestroyProgram(&prog);
    throw ApiError(E_DEVICE, "constraint kernel did not compile: " + log.substr(0, 4000));
  }
  size_t n = 0;
  auto code = std::make_shared<std::string>();
  hiprtcResult rc = hiprtcGetCodeSize(prog, &n);
  if (rc == HIPRTC_SUCCESS) {
    code->resize(n);
    rc = hiprtcGetCode(prog, &(*code)[0]);
  }
  hiprtcDestroyProgram(&prog);
  rtc_check(rc, "hiprtcGetCode");
  std::lock_guard<std::mutex> g(g_air_mu);
  return g_air_code.emplace(std::move(key), std::move(code)).first->second;
}

// The kernel on the current device (loaded on first use; modules stay loaded).
hipFunction_t air_function(int device, const std::shared_ptr<const std::string>& code) {
  std::lock_guard<std::mutex> g(g_air_mu);
  auto key = std::make_pair(device, code.get());
  auto it = g_air_function.find(key);
  if (it != g_air_function.end()) return it->second;
  hipModule_t mod;
  P3G_CHECK_HIP(hipModuleLoadData(&mod, code->data()));
  hipFunction_t fn;
  const hipError_t e = hipModuleGetFunction(&fn, mod, "p3g_air_kernel");
  if (e != hipSuccess) {
    (void)hipModuleUnload(mod);
    P3G_CHECK_HIP(e);
  }
  g_air_function.emplace(key, fn);
  return fn;
}

}  // namespace

uint64_t air_compile(const AirOpIn* ops, uint64_t nops) {
  const AirShape s = analyze(ops, nops);
  return air_code(ops, nops, s)->size();
}

void eval_constraints(Ctx& ctx, const AirOpIn* ops, uint64_t nops, const DevMat& main, const DevMat* aux,
                      const uint64_t* periodic, uint64_t prow, uint64_t pcols, const uint64_t* publics,
                      uint64_t npub, const ef2* ext_inputs, uint64_t next, uint64_t n, unsigned log_d,
                      uint64_t shift, ef2 alpha, ef2* acc, hipStream_t s) {
  if (n == 0 || (n & (n - 1))) throw ApiError(E_NOT_POW2, "trace height must be a power of two");
  const unsigned logn = log2_exact(n);
  if (log_d > AIR_MAX_LOG_D) throw ApiError(E_RANGE, "quotient degree above 2^20");  // Z_H takes d values
  if (logn + log_d > 31) throw ApiError(E_RANGE, "quotient domain above 2^31 points");
  const unsigned L = logn + log_d;
  const uint64_t M = uint64_t(1) << L, d = uint64_t(1) << log_d;
  const AirShape sh = analyze(ops, nops);
  if (!main.ptr || main.height < M) throw ApiError(E_INVAL, "main matrix must hold at least n d rows");
  if (sh.main_cols > main.width) throw ApiError(E_INVAL, "constraint program: main column out of range");
  if (sh.aux) {
    if (!aux || !aux->ptr || aux->height < M || aux->width % 2)
      throw ApiError(E_INVAL, "aux matrix: n d rows of (c0, c1) pairs");
    if (2 * sh.aux_cols > aux->width) throw ApiError(E_INVAL, "constraint program: aux column out of range");
  }
  if (sh.periodic && (!periodic || prow == 0 || prow >= (uint64_t(1) << 32) || sh.per_cols > pcols))
    throw ApiError(E_INVAL, "constraint program: periodic column out of range");
  if (sh.npub > npub) throw ApiError(E_INVAL, "constraint program: public value out of range");
  if (sh.next > next) throw ApiError(E_INVAL, "constraint program: extension input out of range");
  for (uint64_t i = 0; i < sh.npub; i++)
    if (publics[i] >= P) throw ApiError(E_INVAL, "non-canonical public value");
  for (uint64_t i = 0; i < sh.next; i++)
    if (ext_inputs[i].c0 >= P || ext_inputs[i].c1 >= P) throw ApiError(E_INVAL, "non-canonical extension input");
  if (shift >= P) throw ApiError(E_INVAL, "shift is not a canonical field element");
  // Z_H and the selector denominators vanish on the subgroup of order n d
  if (gl_pow(shift, M) == 1) throw ApiError(E_INVAL, "shift must lie outside the subgroup of order n d");

  const auto code = air_code(ops, nops, sh);
  const hipFunction_t fn = air_function(ctx.device, code);

  // one upload: coset tables | alpha powers | publics | extension inputs
  const unsigned k = std::min(AIR_KMAX, L / 2);
  const uint64_t side = uint64_t(1) << k, K = sh.nasserts;
  const uint64_t off_apow = 2 * side + d, off_pub = off_apow + 3 * K, off_ext = off_pub + sh.npub;
  std::vector<uint64_t> tab(off_ext + 2 * sh.next);
  const uint64_t w = gl_two_adic_generator(L);
  const uint64_t w_hi = gl_pow(w, uint64_t(1) << (L - k));
  for (uint64_t l = 0, x = shift, y = 1; l < side; l++, x = gl_mul(x, w), y = gl_mul(y, w_hi)) {
    tab[l] = x;         // shift w^l
    tab[side + l] = y;  // w^(l 2^(L-k))
  }
  const uint64_t shift_n = gl_pow(shift, n), w_d = gl_two_adic_generator(log_d);
  for (uint64_t j = 0, z = shift_n; j < d; j++, z = gl_mul(z, w_d)) tab[2 * side + j] = gl_sub(z, 1);
  ef2 pw = ef_from(1);
  for (uint64_t e = 0; e < K; e++, pw = ef_mul(pw, alpha)) {
    tab[off_apow + 3 * e] = pw.c0;
    tab[off_apow + 3 * e + 1] = pw.c1;
    tab[off_apow + 3 * e + 2] = gl_mul(pw.c1, EXT_W);
  }
  for (uint64_t i = 0; i < sh.npub; i++) tab[off_pub + i] = publics[i];
  for (uint64_t i = 0; i < sh.next; i++) {
    tab[off_ext + 2 * i] = ext_inputs[i].c0;
    tab[off_ext + 2 * i + 1] = ext_inputs[i].c1;
  }

  uint64_t* dtab = (uint64_t*)ctx.upload(tab.data(), tab.size() * sizeof(uint64_t), s);
  uint64_t* dper = nullptr;
  try {
    AirKArgs a{};
    a.main = main.ptr;
    a.main_w = main.width;
    if (sh.aux) {
      a.aux = aux->ptr;
      a.aux_w = aux->width;
    }
    if (sh.periodic) {
      dper = (uint64_t*)ctx.upload(periodic, prow * pcols * sizeof(uint64_t), s);
      a.periodic = dper;
      a.prow = prow;
      a.pcols = pcols;
    }
    a.coset = dtab;
    a.apow = dtab + off_apow;
    a.pub = dtab + off_pub;
    a.ext = reinterpret_cast<const ef2*>(dtab + off_ext);
    a.acc = acc;
    a.omega_tile = gl_pow(w, side);
    a.omega_n_inv = gl_inv(gl_two_adic_generator(logn));
    a.logM = L;
    a.k = k;
    a.log_d = log_d;
    a.blocks_on = ~0u;
    void* params[] = {&a};
    const uint32_t blocks = (uint32_t)(M >> (2 * k));
    ProfScope ps(ctx, "eval_constraints", s);
    P3G_CHECK_HIP(hipModuleLaunchKernel(fn, blocks, 1, 1, AIR_T, 1, 1, 0, s, params, nullptr));
  } catch (...) {
    ctx.release(dtab);
    ctx.release(dper);
    throw;
  }
  ctx.release(dtab);
  ctx.release(dper);
}

}  // namespace p3g
