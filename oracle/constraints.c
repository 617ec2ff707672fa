/*
 * oracle/constraints.c — constraint evaluation on the quotient domain, plain C + OpenMP.
 *
 * TEST INFRASTRUCTURE ONLY (the checker; never linked by product code).
 *
 * Restates evaluate_constraints_into (p3-miden-lifted-stark/src/prover/constraints/
 * mod.rs:63-179) for an AIR given as a straight-line program (the op list of
 * include/p3gpu.h, one value per op, as a symbolic AIR builder records it):
 *   - the quotient domain gJ = shift * <w_M>, M = n d, natural order i; main and aux
 *     come as the committed LDEs with bit-reversed rows, whose first M rows are gJ, so
 *     point i is stored row bitrev_M(i) (Committed::evals_on_quotient_domain,
 *     commit.rs:134-144); the next row is point (i + d) mod M, the step of
 *     vertically_packed_row_pair(i, constraint_degree) (mod.rs:123-131);
 *   - aux extension columns as (c0, c1) pairs of base columns (mod.rs:133-143);
 *   - periodic values from the table's row i mod rows (PeriodicLde::packed_values_at,
 *     periodic.rs:86-99);
 *   - selectors unnormalised: Z_H(x)/(x - 1), Z_H(x)/(x - w_n^-1), x - w_n^-1,
 *     Z_H(x) = x^n - 1 (coset.rs:210-247);
 *   - the folded value sum_k alpha^(K-1-k) C_k over the constraints in assertion order,
 *     with the powers of alpha computed explicitly (the reference splits them into base
 *     and extension coefficient vectors, layout.rs / folder.rs:82-97; the sum is the same);
 *   - acc[i] += folded (mod.rs:170-175).
 */
#include <stdlib.h>
#include <string.h>
#include <omp.h>
#include "oracle.h"

enum { A_CONST, A_MAIN, A_AUX, A_PERIODIC, A_PUBLIC, A_EXT, A_FIRST, A_LAST, A_TRANS,
       A_ADD, A_SUB, A_MUL, A_NEG, A_ASSERT, A_ASSERT_EXT };

int orc_eval_constraints(const orc_air_op* ops, size_t nops, const uint64_t* main, size_t main_h, size_t main_w,
                         const uint64_t* aux, size_t aux_h, size_t aux_w, const uint64_t* periodic, size_t prow,
                         size_t pcols, const uint64_t* publics, size_t npub, const uint64_t* ext, size_t next,
                         size_t n, unsigned log_d, uint64_t shift, const uint64_t* alpha, uint64_t* acc) {
    if (n == 0 || (n & (n - 1)) || nops == 0) return -1;
    const unsigned logn = log2_exact(n), logM = logn + log_d;
    const size_t M = (size_t)1 << logM, d = (size_t)1 << log_d;
    if (main_h < M || (aux && aux_h < M)) return -1;
    /* the coset must miss the subgroup of order M, where Z_H and the selector
       denominators vanish */
    if (shift >= GL_P || gl_pow(shift, M) == 1) return -1;
    /* check the program and count the constraints; operands name earlier values */
    size_t K = 0;
    unsigned char* is_assert = (unsigned char*)calloc(nops, 1);
#define OPERAND(v) if ((v) >= k || is_assert[(v)]) { free(is_assert); return -1; }
    for (size_t k = 0; k < nops; k++) {
        const orc_air_op* o = &ops[k];
        is_assert[k] = o->op == A_ASSERT || o->op == A_ASSERT_EXT;
        switch (o->op) {
            case A_MAIN: if (o->a >= main_w || o->b > 1) { free(is_assert); return -1; } break;
            case A_AUX: if (!aux || 2 * (size_t)o->a + 1 >= aux_w || o->b > 1) { free(is_assert); return -1; } break;
            case A_PERIODIC: if (!periodic || o->a >= pcols || prow == 0) { free(is_assert); return -1; } break;
            case A_PUBLIC: if (o->a >= npub) { free(is_assert); return -1; } break;
            case A_EXT: if (o->a >= next) { free(is_assert); return -1; } break;
            case A_ADD: case A_SUB: case A_MUL: OPERAND(o->a) OPERAND(o->b) break;
            case A_NEG: OPERAND(o->a) break;
            case A_ASSERT: case A_ASSERT_EXT: OPERAND(o->a) K++; break;
            case A_CONST: if (o->imm >= GL_P) { free(is_assert); return -1; } break;
            case A_FIRST: case A_LAST: case A_TRANS: break;
            default: free(is_assert); return -1;
        }
    }
#undef OPERAND
    free(is_assert);
    if (K == 0) return -1;
    /* alpha^(K-1-k) for the k-th constraint */
    const ef2 al = ef_make(alpha[0], alpha[1]);
    ef2* pw = (ef2*)malloc(K * sizeof(ef2));
    pw[K - 1] = ef_from(1);
    for (size_t k = K - 1; k-- > 0;) pw[k] = ef_mul(pw[k + 1], al);
    const gl om_M = gl_two_adic_generator(logM);
    const gl om_n_inv = gl_inv(gl_two_adic_generator(logn));
    #pragma omp parallel
    {
        ef2* val = (ef2*)malloc(nops * sizeof(ef2));
        #pragma omp for schedule(static)
        for (size_t i = 0; i < M; i++) {
            const size_t cur = bitrev_idx(i, logM), nxt = bitrev_idx((i + d) % M, logM);
            const gl x = gl_mul(shift, gl_pow(om_M, i));
            const gl zh = gl_sub(gl_pow(x, n), 1);
            const ef2 first = ef_from(gl_mul(zh, gl_inv(gl_sub(x, 1))));
            const ef2 last = ef_from(gl_mul(zh, gl_inv(gl_sub(x, om_n_inv))));
            const ef2 trans = ef_from(gl_sub(x, om_n_inv));
            ef2 folded = ef_from(0);
            size_t kc = 0;
            for (size_t k = 0; k < nops; k++) {
                const orc_air_op* o = &ops[k];
                ef2 v = ef_from(0);
                switch (o->op) {
                    case A_CONST: v = ef_from(o->imm); break;
                    case A_MAIN: v = ef_from(main[(o->b ? nxt : cur) * main_w + o->a]); break;
                    case A_AUX: {
                        const uint64_t* r = aux + (o->b ? nxt : cur) * aux_w + 2 * (size_t)o->a;
                        v = ef_make(r[0], r[1]);
                        break;
                    }
                    case A_PERIODIC: v = ef_from(periodic[(i % prow) * pcols + o->a]); break;
                    case A_PUBLIC: v = ef_from(publics[o->a]); break;
                    case A_EXT: v = ef_make(ext[2 * (size_t)o->a], ext[2 * (size_t)o->a + 1]); break;
                    case A_FIRST: v = first; break;
                    case A_LAST: v = last; break;
                    case A_TRANS: v = trans; break;
                    case A_ADD: v = ef_add(val[o->a], val[o->b]); break;
                    case A_SUB: v = ef_sub(val[o->a], val[o->b]); break;
                    case A_MUL: v = ef_mul(val[o->a], val[o->b]); break;
                    case A_NEG: v = ef_neg(val[o->a]); break;
                    default: folded = ef_add(folded, ef_mul(pw[kc++], val[o->a])); break;
                }
                val[k] = v;
            }
            ef2 a = ef_make(acc[2 * i], acc[2 * i + 1]);
            a = ef_add(a, folded);
            acc[2 * i] = a.c0;
            acc[2 * i + 1] = a.c1;
        }
        free(val);
    }
    free(pw);
    return 0;
}
