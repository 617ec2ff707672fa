/* oracle/tx_internal.h — transcript layout shared by the oracle's C files (test infrastructure). */
#ifndef ORACLE_TX_INTERNAL_H
#define ORACLE_TX_INTERNAL_H
#include "oracle.h"

typedef struct {
    int kind;
    /* DuplexChallenger<F, Poseidon2, 12, 8> */
    gl st[12];
    gl in[8];
    int in_len;
    gl out[8];
    int out_len;
    /* SerializingChallenger64<F, HashChallenger<u8, Keccak256, 32>> */
    uint8_t* kin;
    size_t kin_len, kin_cap;
    uint8_t kout[32];
    int kout_len;
} orc_challenger;

struct orc_tx {
    orc_challenger ch;
    uint64_t* fields;
    size_t nf, capf;
    uint64_t* comms;
    size_t nc, capc;
    size_t rf, rc; /* verifier read cursors */
};

#endif
