/*
 * sbf_oracle.h -- CPU restatement of the Super Bloom k-mer insert/query path.
 *
 * TEST INFRASTRUCTURE ONLY. This oracle is the checker that the HIP product
 * path is compared against. Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may load it; the product library never links
 * or calls it.
 *
 * Two forms are provided:
 *   - the naive form (orc_insert / orc_contains ...) recomputes every hash per
 *     k-mer window exactly as the reference kernels do, and is the parity
 *     oracle;
 *   - the fast form (orc_mt_insert / orc_mt_contains) computes each m-mer and
 *     s-mer hash once per position with sliding windows on all host threads;
 *     it is cross-checked against the naive form and timed as the CPU baseline.
 *
 * Reference files restated (paths relative to the reference checkout):
 *   include/cusbf/config.cuh:28-118, include/cusbf/Alphabet.cuh:113-218,
 *   include/cusbf/hashutil.cuh:192-209, include/cusbf/detail/filter_common.cuh:22-139,
 *   include/cusbf/detail/filter_impl.cuh:18-77, include/cusbf/filter_ref.cuh:52-108,
 *   include/cusbf/detail/sequence_kmer.cuh:21-156, include/cusbf/detail/kernels.cuh:40-172,
 *   include/cusbf/detail/dense_packed.cuh:13-149, include/cusbf/dense_packed.hpp:35-49,
 *   include/cusbf/filter.cuh:108-120.
 */
#ifndef SBF_ORACLE_H
#define SBF_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
    ORC_ALPHA_DNA = 0,         /* Alphabet.cuh:113-142 */
    ORC_ALPHA_DNA_TRIPLET = 1, /* Alphabet.cuh:150-178 */
    ORC_ALPHA_PROTEIN = 2,     /* Alphabet.cuh:189-218 */
    ORC_ALPHA_CUSTOM = 3       /* width-1 alphabet given as a byte->symbol table */
};

typedef struct orc_config {
    uint32_t k, s, m, hash_count;
    uint32_t alphabet;
    uint32_t symbol_width;
    uint32_t symbol_count;
    uint32_t symbol_bits;
    uint8_t invalid_symbol;
    uint8_t separator;
    uint8_t custom_lut[256];
} orc_config;

/* Returns 0 on success, -1 if the parameters violate config.cuh:74-117. */
int orc_config_init(orc_config* c, uint32_t k, uint32_t s, uint32_t m, uint32_t hash_count,
                    uint32_t alphabet, const uint8_t* custom_lut, uint32_t custom_symbol_count,
                    uint8_t custom_invalid, uint8_t custom_separator);

uint8_t orc_encode(const orc_config* c, const char* p);
uint64_t orc_hash64(uint64_t key);
uint64_t orc_minimizer_hash64(uint64_t key);
uint64_t orc_salt(uint32_t index);
uint32_t orc_bit_address(const orc_config* c, uint32_t hash_index, uint64_t base_hash);
uint64_t orc_num_shards(uint64_t requested_filter_bits);
uint64_t orc_kmer_count(const orc_config* c, uint64_t bytes);

/* Per-window helpers of the naive form. */
int orc_window_valid(const orc_config* c, const char* seq, uint64_t window);
uint64_t orc_pack_window(const orc_config* c, const char* seq, uint64_t window);
uint64_t orc_window_minimizer_hash(const orc_config* c, uint64_t packed_kmer);
uint32_t orc_shard_index(uint64_t minimizer_hash, uint64_t num_shards);
void orc_window_masks(const orc_config* c, uint64_t packed_kmer, uint64_t masks[4]);

/* Naive byte path. `words` holds num_shards*4 uint64 filter words. */
void orc_insert(const orc_config* c, const char* seq, uint64_t bytes, uint64_t* words,
                uint64_t num_shards);
void orc_contains(const orc_config* c, const char* seq, uint64_t bytes, const uint64_t* words,
                  uint64_t num_shards, uint8_t* out);

/* Dense packed path (LSB-first, 64/symbol_bits symbols per word). */
uint64_t orc_dense_word_count(const orc_config* c, uint64_t num_symbols);
void orc_pack_dense(const orc_config* c, const char* seq, uint64_t num_symbols, uint64_t* out);
void orc_insert_dense(const orc_config* c, const uint64_t* packed, uint64_t num_symbols,
                      uint64_t* words, uint64_t num_shards);
void orc_contains_dense(const orc_config* c, const uint64_t* packed, uint64_t num_symbols,
                        const uint64_t* words, uint64_t num_shards, uint8_t* out);

/* Counts over the byte path: valid windows and shard runs (maximal runs of
 * consecutive valid windows sharing a shard index). */
void orc_count(const orc_config* c, const char* seq, uint64_t bytes, uint64_t num_shards,
               uint64_t* valid_windows, uint64_t* shard_runs);

/* Fast multithreaded form (CPU Super Bloom baseline). threads<=0 -> all cores. */
void orc_mt_insert(const orc_config* c, const char* seq, uint64_t bytes, uint64_t* words,
                   uint64_t num_shards, int threads);
void orc_mt_contains(const orc_config* c, const char* seq, uint64_t bytes, const uint64_t* words,
                     uint64_t num_shards, uint8_t* out, int threads);
int orc_hardware_threads(void);

#ifdef __cplusplus
}
#endif
#endif
