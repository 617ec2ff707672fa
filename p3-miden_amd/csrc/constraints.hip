// Constraint evaluation on the quotient domain, for gfx950.
//
// Replaces evaluate_constraints_into (p3-miden-lifted-stark/src/prover/constraints/
// mod.rs:63-179): at every point x_i = shift w^i of the quotient domain gJ
// (natural order, |gJ| = n d), evaluate the AIR's constraints on the current and
// next trace rows, fold them as sum_k alpha^(K-1-k) C_k, and add the result to
// the accumulator.
//
// The AIR arrives as a straight-line program (include/p3gpu.h, p3g_air_op), the
// form a symbolic AIR builder records.  The host checks it, folds public values
// and extension inputs into constants, and assigns its values to at most 64
// registers by liveness (a value's register is reused once its last reader has
// run).  Each thread then interprets the program for one point: the program sits
// in LDS, the registers in the thread's private memory.
//
// Rows are read in place from the committed LDE: its first n d rows, in
// bit-reversed order, are gJ (commit.rs:134-144), so point i is stored row
// bitrev(i) and its next row is point (i + d) mod n d.
#include <string>

#include "internal.hpp"

namespace p3g {

namespace {

constexpr int AIR_T = 256;
constexpr uint32_t AIR_MAX_INST = 1024;
constexpr uint32_t AIR_REGS = 64;

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

// One instruction after register assignment.
struct AirInst {
  uint32_t op;
  uint32_t dst, ra, rb;  // registers
  uint32_t col, row;     // MAIN / AUX / PERIODIC operands
  uint64_t c0, c1;       // CONST (public values and extension inputs folded in)
};

struct AirArgs {
  const AirInst* prog;
  uint32_t ninst;
  uint32_t need_sel;
  const uint64_t* main;
  uint64_t main_w;
  const uint64_t* aux;
  uint64_t aux_w;  // base words per row
  const uint64_t* periodic;
  uint64_t prow, pcols;
  uint64_t M;  // n d
  uint32_t logM;
  uint64_t d;
  uint64_t shift, omega_M, omega_d, shift_pow_n, omega_n_inv;
  ef2 alpha;
  ef2* acc;
};

__global__ void __launch_bounds__(AIR_T) air_eval_kernel(AirArgs a) {
  __shared__ AirInst prog[AIR_MAX_INST];
  for (uint32_t k = threadIdx.x; k < a.ninst; k += blockDim.x) prog[k] = a.prog[k];
  __syncthreads();
  const uint64_t i = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= a.M) return;
  const uint64_t cur = bitrev32((uint32_t)i, a.logM);
  const uint64_t nxt = bitrev32((uint32_t)((i + a.d) & (a.M - 1)), a.logM);
  ef2 first{0, 0}, last{0, 0}, trans{0, 0};
  if (a.need_sel) {
    const uint64_t x = gl_mul(a.shift, gl_pow(a.omega_M, i));
    // x^n = shift^n w_d^i, since w_M^n = w_d
    const uint64_t zh = gl_sub(gl_mul(a.shift_pow_n, gl_pow(a.omega_d, i & (a.d - 1))), 1);
    const uint64_t xm = gl_sub(x, a.omega_n_inv);
    first = ef_from(gl_mul(zh, gl_inv(gl_sub(x, 1))));
    last = ef_from(gl_mul(zh, gl_inv(xm)));
    trans = ef_from(xm);
  }
  ef2 reg[AIR_REGS];
  ef2 folded{0, 0};
  for (uint32_t k = 0; k < a.ninst; k++) {
    const AirInst& in = prog[k];
    switch (in.op) {
      case OP_CONST: reg[in.dst] = ef2{in.c0, in.c1}; break;
      case OP_MAIN: reg[in.dst] = ef_from(a.main[(in.row ? nxt : cur) * a.main_w + in.col]); break;
      case OP_AUX: {
        const uint64_t* r = a.aux + (in.row ? nxt : cur) * a.aux_w + 2 * in.col;
        reg[in.dst] = ef2{r[0], r[1]};
        break;
      }
      case OP_PERIODIC: reg[in.dst] = ef_from(a.periodic[(i % a.prow) * a.pcols + in.col]); break;
      case OP_FIRST: reg[in.dst] = first; break;
      case OP_LAST: reg[in.dst] = last; break;
      case OP_TRANSITION: reg[in.dst] = trans; break;
      case OP_ADD: reg[in.dst] = ef_add(reg[in.ra], reg[in.rb]); break;
      case OP_SUB: reg[in.dst] = ef_sub(reg[in.ra], reg[in.rb]); break;
      case OP_MUL: reg[in.dst] = ef_mul(reg[in.ra], reg[in.rb]); break;
      case OP_NEG: reg[in.dst] = ef_neg(reg[in.ra]); break;
      default: folded = ef_add(ef_mul(folded, a.alpha), reg[in.ra]); break;  // an assert (Horner in alpha)
    }
  }
  a.acc[i] = ef_add(a.acc[i], folded);
}

}  // namespace

void eval_constraints(Ctx& ctx, const AirOpIn* ops, uint64_t nops, const DevMat& main, const DevMat* aux,
                      const uint64_t* periodic, uint64_t prow, uint64_t pcols, const uint64_t* publics,
                      uint64_t npub, const ef2* ext_inputs, uint64_t next, uint64_t n, unsigned log_d,
                      uint64_t shift, ef2 alpha, ef2* acc, hipStream_t s) {
  if (n == 0 || (n & (n - 1))) throw ApiError(E_NOT_POW2, "trace height must be a power of two");
  const unsigned logn = log2_exact(n);
  const unsigned logM = logn + log_d;
  if (logM > 31) throw ApiError(E_RANGE, "quotient domain above 2^31 points");
  const uint64_t M = uint64_t(1) << logM;
  if (nops == 0) throw ApiError(E_EMPTY, "empty constraint program");
  if (nops > AIR_MAX_INST) throw ApiError(E_RANGE, "constraint program longer than 1024 ops");
  if (!main.ptr || main.height < M) throw ApiError(E_INVAL, "main matrix must hold at least n d rows");
  if (aux && aux->ptr && (aux->height < M || aux->width % 2)) throw ApiError(E_INVAL, "aux matrix: n d rows of (c0, c1) pairs");
  if (shift >= P) throw ApiError(E_INVAL, "shift is not a canonical field element");

  // check the program, fold inputs into constants, find each value's last reader
  std::vector<int64_t> last_use(nops, -1);
  std::vector<bool> is_value(nops, false);
  bool need_sel = false;
  uint32_t nasserts = 0;
  for (uint64_t k = 0; k < nops; k++) {
    const auto& o = ops[k];
    auto operand = [&](uint32_t v) {
      if (v >= k || !is_value[v]) throw ApiError(E_INVAL, "constraint program: op " + std::to_string(k) + " reads an undefined value");
      last_use[v] = (int64_t)k;
    };
    switch (o.op) {
      case OP_CONST:
        if (o.imm >= P) throw ApiError(E_INVAL, "constraint program: non-canonical constant");
        break;
      case OP_MAIN:
        if (o.a >= main.width || o.b > 1) throw ApiError(E_INVAL, "constraint program: main column or row out of range");
        break;
      case OP_AUX:
        if (!aux || !aux->ptr || 2ull * o.a + 1 >= aux->width || o.b > 1)
          throw ApiError(E_INVAL, "constraint program: aux column or row out of range");
        break;
      case OP_PERIODIC:
        if (!periodic || prow == 0 || o.a >= pcols)
          throw ApiError(E_INVAL, "constraint program: periodic column out of range");
        break;
      case OP_PUBLIC:
        if (o.a >= npub) throw ApiError(E_INVAL, "constraint program: public value out of range");
        break;
      case OP_EXT_INPUT:
        if (o.a >= next) throw ApiError(E_INVAL, "constraint program: extension input out of range");
        break;
      case OP_FIRST:
      case OP_LAST:
      case OP_TRANSITION: need_sel = true; break;
      case OP_ADD:
      case OP_SUB:
      case OP_MUL:
        operand(o.a);
        operand(o.b);
        break;
      case OP_NEG: operand(o.a); break;
      case OP_ASSERT:
      case OP_ASSERT_EXT:
        operand(o.a);
        nasserts++;
        break;
      default: throw ApiError(E_INVAL, "constraint program: unknown op " + std::to_string(o.op));
    }
    is_value[k] = o.op != OP_ASSERT && o.op != OP_ASSERT_EXT;
  }
  if (nasserts == 0) throw ApiError(E_EMPTY, "constraint program asserts nothing");

  // registers by liveness: a value's register is free after its last reader
  std::vector<AirInst> prog;
  prog.reserve(nops);
  std::vector<uint32_t> reg_of(nops, 0);
  std::vector<uint32_t> free_regs;
  for (uint32_t r = AIR_REGS; r-- > 0;) free_regs.push_back(r);
  for (uint64_t k = 0; k < nops; k++) {
    const auto& o = ops[k];
    AirInst in{};
    in.op = o.op;
    in.col = o.a;
    in.row = o.b;
    if (o.op == OP_ADD || o.op == OP_SUB || o.op == OP_MUL) {
      in.ra = reg_of[o.a];
      in.rb = reg_of[o.b];
    } else if (o.op == OP_NEG || o.op == OP_ASSERT || o.op == OP_ASSERT_EXT) {
      in.ra = reg_of[o.a];
    }
    if (o.op == OP_CONST) {
      in.c0 = o.imm;
    } else if (o.op == OP_PUBLIC) {
      in.op = OP_CONST;
      if (publics[o.a] >= P) throw ApiError(E_INVAL, "non-canonical public value");
      in.c0 = publics[o.a];
    } else if (o.op == OP_EXT_INPUT) {
      in.op = OP_CONST;
      in.c0 = ext_inputs[o.a].c0;
      in.c1 = ext_inputs[o.a].c1;
      if (in.c0 >= P || in.c1 >= P) throw ApiError(E_INVAL, "non-canonical extension input");
    }
    // operands whose last reader is this op give their registers back first, so the
    // result may take one of them (every op reads its operands before it writes)
    auto release = [&](uint32_t v) {
      if (last_use[v] == (int64_t)k) {
        free_regs.push_back(reg_of[v]);
        last_use[v] = -2;  // released once even when both operands are the same value
      }
    };
    if (o.op == OP_ADD || o.op == OP_SUB || o.op == OP_MUL) {
      release(o.a);
      release(o.b);
    } else if (o.op == OP_NEG || o.op == OP_ASSERT || o.op == OP_ASSERT_EXT) {
      release(o.a);
    }
    if (is_value[k]) {
      if (last_use[k] == -1) continue;  // never read: no instruction needed
      if (free_regs.empty()) throw ApiError(E_RANGE, "constraint program keeps more than 64 values live");
      reg_of[k] = free_regs.back();
      free_regs.pop_back();
      in.dst = reg_of[k];
    }
    prog.push_back(in);
  }

  AirArgs a{};
  a.ninst = (uint32_t)prog.size();
  a.need_sel = need_sel;
  a.main = main.ptr;
  a.main_w = main.width;
  a.aux = aux ? aux->ptr : nullptr;
  a.aux_w = aux ? aux->width : 0;
  a.prow = prow;
  a.pcols = pcols;
  a.M = M;
  a.logM = logM;
  a.d = uint64_t(1) << log_d;
  a.shift = shift;
  a.omega_M = gl_two_adic_generator(logM);
  a.omega_d = gl_two_adic_generator(log_d);
  a.shift_pow_n = gl_pow(shift, n);
  a.omega_n_inv = gl_inv(gl_two_adic_generator(logn));
  a.alpha = alpha;
  a.acc = acc;
  AirInst* dprog = (AirInst*)ctx.upload(prog.data(), prog.size() * sizeof(AirInst), s);
  uint64_t* dper = nullptr;
  try {
    a.prog = dprog;
    if (periodic && prow && pcols) {
      dper = (uint64_t*)ctx.upload(periodic, prow * pcols * sizeof(uint64_t), s);
      a.periodic = dper;
    }
    ProfScope ps(ctx, "eval_constraints", s);
    hipLaunchKernelGGL(air_eval_kernel, dim3((uint32_t)((M + AIR_T - 1) / AIR_T)), dim3(AIR_T), 0, s, a);
    P3G_CHECK_HIP(hipGetLastError());
  } catch (...) {
    ctx.release(dprog);
    ctx.release(dper);
    throw;
  }
  ctx.release(dprog);
  ctx.release(dper);
}

}  // namespace p3g
