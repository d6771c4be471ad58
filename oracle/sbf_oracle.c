/*
 * sbf_oracle.c -- CPU restatement of the Super Bloom insert/query path.
 *
 * TEST INFRASTRUCTURE ONLY (the checker, never the product): see sbf_oracle.h.
 * Every function cites the reference file:line it restates; paths are relative
 * to the reference checkout.
 */
#define _GNU_SOURCE
#include "sbf_oracle.h"

#include <pthread.h>
#include <stdlib.h>
#include <string.h>
#include <unistd.h>

/* hashutil.cuh:192-196 (hash64) and :207-209 (minimizer_hash64) share this multiplier. */
#define ORC_GOLDEN 0x9E3779B97F4A7C15ULL

/* filter_common.cuh:27-93: the 16 multiplicative salts (SaltLiteral<0..15>). */
static const uint64_t kSalts[16] = {
    0x9E3779B97F4A7C15ULL, 0xC2B2AE3D27D4EB4FULL, 0x165667B19E3779F9ULL, 0x85EBCA77C2B2AE63ULL,
    0x27D4EB2F165667C5ULL, 0x94D049BB133111EFULL, 0xBF58476D1CE4E5B9ULL, 0xD6E8FEB86659FD93ULL,
    0xA0761D6478BD642FULL, 0xE7037ED1A0B428DBULL, 0x8EBC6AF09C88C6E3ULL, 0x589965CC75374CC3ULL,
    0x1D8E4E27C47D124FULL, 0xEB449C93FBBEA6B5ULL, 0xDB4F0B9175AE2165ULL, 0xBBE056FDADE14B91ULL,
};

static uint32_t bit_width_u32(uint32_t x) {
    uint32_t w = 0;
    while (x) {
        ++w;
        x >>= 1;
    }
    return w;
}

/* config.cuh:28-118: derived constants and the static_assert set. */
int orc_config_init(orc_config* c, uint32_t k, uint32_t s, uint32_t m, uint32_t hash_count,
                    uint32_t alphabet, const uint8_t* custom_lut, uint32_t custom_symbol_count,
                    uint8_t custom_invalid, uint8_t custom_separator) {
    memset(c, 0, sizeof(*c));
    c->k = k;
    c->s = s;
    c->m = m;
    c->hash_count = hash_count;
    c->alphabet = alphabet;
    switch (alphabet) {
        case ORC_ALPHA_DNA: /* Alphabet.cuh:113-125 */
            c->symbol_width = 1;
            c->symbol_count = 4;
            c->invalid_symbol = 0xFF;
            c->separator = 'N';
            break;
        case ORC_ALPHA_DNA_TRIPLET: /* Alphabet.cuh:150-162 */
            c->symbol_width = 3;
            c->symbol_count = 64;
            c->invalid_symbol = 0xFF;
            c->separator = 'N';
            break;
        case ORC_ALPHA_PROTEIN: /* Alphabet.cuh:189-201 */
            c->symbol_width = 1;
            c->symbol_count = 26;
            c->invalid_symbol = 0xFF;
            c->separator = '*';
            break;
        case ORC_ALPHA_CUSTOM:
            if (!custom_lut || custom_symbol_count == 0 || custom_symbol_count > 255) return -1;
            c->symbol_width = 1;
            c->symbol_count = custom_symbol_count;
            c->invalid_symbol = custom_invalid;
            c->separator = custom_separator;
            memcpy(c->custom_lut, custom_lut, 256);
            break;
        default:
            return -1;
    }
    c->symbol_bits = bit_width_u32(c->symbol_count - 1); /* config.cuh:51 */
    if (c->symbol_bits == 0) return -1;
    if (k == 0 || m == 0 || m > k || s == 0 || s > k) return -1; /* config.cuh:74-77 */
    if (k * c->symbol_bits > 64) return -1;                      /* config.cuh:78 */
    if (hash_count < 4 || hash_count > 16 || hash_count % 4 != 0) return -1; /* :81-82,104-111 */
    return 0;
}

/* Alphabet.cuh:132-141 (DNA), :167-177 (triplet), :208-217 (protein). */
static uint8_t encode_dna(uint8_t byte) {
    const uint8_t upper = byte & 0xDFu;
    const int valid = upper == 'A' || upper == 'C' || upper == 'G' || upper == 'T';
    return valid ? (uint8_t)((byte >> 1u) & 3u) : 0xFFu;
}

uint8_t orc_encode(const orc_config* c, const char* p) {
    const uint8_t b0 = (uint8_t)p[0];
    switch (c->alphabet) {
        case ORC_ALPHA_DNA:
            return encode_dna(b0);
        case ORC_ALPHA_DNA_TRIPLET: {
            const uint8_t a = encode_dna(b0);
            const uint8_t b = encode_dna((uint8_t)p[1]);
            const uint8_t d = encode_dna((uint8_t)p[2]);
            if (a == 0xFF || b == 0xFF || d == 0xFF) return 0xFF;
            return (uint8_t)((a << 4u) | (b << 2u) | d);
        }
        case ORC_ALPHA_PROTEIN: {
            const uint8_t idx = (uint8_t)((b0 & 0xDFu) - 'A');
            return idx < 26 ? idx : 0xFF;
        }
        default:
            return c->custom_lut[b0];
    }
}

/* hashutil.cuh:192-196 */
uint64_t orc_hash64(uint64_t key) {
    key *= ORC_GOLDEN;
    key ^= key >> 33;
    return key;
}

/* hashutil.cuh:207-209 */
uint64_t orc_minimizer_hash64(uint64_t key) { return key * ORC_GOLDEN; }

uint64_t orc_salt(uint32_t index) { return kSalts[index & 15u]; }

/* filter_impl.cuh:18-47: bit-slicing when 64/h >= 6, else salted multiply-shift. */
uint32_t orc_bit_address(const orc_config* c, uint32_t hash_index, uint64_t base_hash) {
    const uint32_t slice_width = 64u / c->hash_count;
    if (slice_width >= 6u) {
        return (uint32_t)((base_hash >> (slice_width * hash_index)) & 63u);
    }
    return (uint32_t)((base_hash * kSalts[hash_index]) >> 58);
}

/* filter.cuh:108-117: bit_ceil(max(1, ceil(bits / 256))). */
uint64_t orc_num_shards(uint64_t requested_filter_bits) {
    uint64_t shards = requested_filter_bits / 256u + (requested_filter_bits % 256u != 0);
    if (shards < 1) shards = 1;
    uint64_t p = 1;
    while (p < shards) p <<= 1;
    return p;
}

/* record_math.cuh:12-16, sequence_kmer.cuh:27-30 */
uint64_t orc_kmer_count(const orc_config* c, uint64_t bytes) {
    const uint64_t symbols = bytes / c->symbol_width;
    return symbols < c->k ? 0 : symbols - c->k + 1;
}

/* filter_common.cuh:118-125 */
static uint64_t window_mask(const orc_config* c, uint32_t len) {
    const uint32_t bits = c->symbol_bits * len;
    return bits >= 64 ? ~0ULL : ((1ULL << bits) - 1);
}

/* filter_common.cuh:128-134 (MSB-first sub-window of a packed k-mer). */
static uint64_t extract_subwindow(const orc_config* c, uint64_t packed, uint32_t len,
                                  uint32_t start) {
    return (packed >> (c->symbol_bits * (c->k - (start + len)))) & window_mask(c, len);
}

static uint8_t symbol_at(const orc_config* c, const char* seq, uint64_t index) {
    return orc_encode(c, seq + index * c->symbol_width);
}

/* sequence_kmer.cuh:92-101 */
int orc_window_valid(const orc_config* c, const char* seq, uint64_t window) {
    for (uint32_t i = 0; i < c->k; ++i) {
        if (symbol_at(c, seq, window + i) == c->invalid_symbol) return 0;
    }
    return 1;
}

/* sequence_kmer.cuh:74-82 */
uint64_t orc_pack_window(const orc_config* c, const char* seq, uint64_t window) {
    const uint64_t symbol_mask = (1ULL << c->symbol_bits) - 1;
    uint64_t packed = 0;
    for (uint32_t i = 0; i < c->k; ++i) {
        packed = (packed << c->symbol_bits) | (symbol_at(c, seq, window + i) & symbol_mask);
    }
    return packed;
}

/* sequence_kmer.cuh:40-50 */
uint64_t orc_window_minimizer_hash(const orc_config* c, uint64_t packed_kmer) {
    uint64_t best = ~0ULL; /* kInvalidHash, filter_common.cuh:25 */
    for (uint32_t o = 0; o + c->m <= c->k; ++o) {
        const uint64_t h = orc_minimizer_hash64(extract_subwindow(c, packed_kmer, c->m, o));
        if (h < best) best = h;
    }
    return best;
}

/* filter_ref.cuh:52-55 with the uint32_t narrowing of kernels.cuh:240-241 / :307-308. */
uint32_t orc_shard_index(uint64_t minimizer_hash, uint64_t num_shards) {
    return (uint32_t)(minimizer_hash & (num_shards - 1));
}

/* kernels.cuh:294-304 with filter_impl.cuh:54-77 (s-mer hashes OR'ed into 4 word masks). */
void orc_window_masks(const orc_config* c, uint64_t packed_kmer, uint64_t masks[4]) {
    masks[0] = masks[1] = masks[2] = masks[3] = 0;
    for (uint32_t o = 0; o + c->s <= c->k; ++o) {
        const uint64_t h = orc_hash64(extract_subwindow(c, packed_kmer, c->s, o));
        for (uint32_t i = 0; i < c->hash_count; ++i) {
            masks[i % 4u] |= 1ULL << orc_bit_address(c, i, h);
        }
    }
}

/* filter_ref.cuh:70-87 */
static int contains_packed(const orc_config* c, uint64_t packed_kmer, const uint64_t* shard) {
    for (uint32_t o = 0; o + c->s <= c->k; ++o) {
        const uint64_t h = orc_hash64(extract_subwindow(c, packed_kmer, c->s, o));
        for (uint32_t i = 0; i < c->hash_count; ++i) {
            if (((shard[i % 4u] >> orc_bit_address(c, i, h)) & 1u) == 0) return 0;
        }
    }
    return 1;
}

/* kernels.cuh:106-172 (insert_kmers_from_symbol_tile): every valid window ORs its
 * masks into its shard; filter_ref.cuh:94-108 applies them. The warp segmented
 * reduction only merges ORs, so the resulting words equal this per-window loop. */
void orc_insert(const orc_config* c, const char* seq, uint64_t bytes, uint64_t* words,
                uint64_t num_shards) {
    const uint64_t n = orc_kmer_count(c, bytes);
    for (uint64_t j = 0; j < n; ++j) {
        if (!orc_window_valid(c, seq, j)) continue;
        const uint64_t packed = orc_pack_window(c, seq, j);
        const uint32_t shard = orc_shard_index(orc_window_minimizer_hash(c, packed), num_shards);
        uint64_t masks[4];
        orc_window_masks(c, packed, masks);
        for (int w = 0; w < 4; ++w) words[(uint64_t)shard * 4u + w] |= masks[w];
    }
}

/* kernels.cuh:40-99 (contains_kmers_from_symbol_tile): invalid -> 0, else findere test. */
void orc_contains(const orc_config* c, const char* seq, uint64_t bytes, const uint64_t* words,
                  uint64_t num_shards, uint8_t* out) {
    const uint64_t n = orc_kmer_count(c, bytes);
    for (uint64_t j = 0; j < n; ++j) {
        if (!orc_window_valid(c, seq, j)) {
            out[j] = 0;
            continue;
        }
        const uint64_t packed = orc_pack_window(c, seq, j);
        const uint32_t shard = orc_shard_index(orc_window_minimizer_hash(c, packed), num_shards);
        out[j] = (uint8_t)contains_packed(c, packed, words + (uint64_t)shard * 4u);
    }
}

/* dense_packed.cuh:13-22 */
uint64_t orc_dense_word_count(const orc_config* c, uint64_t num_symbols) {
    const uint64_t per_word = 64u / c->symbol_bits;
    return (num_symbols + per_word - 1) / per_word;
}

/* dense_packed.hpp:35-49: LSB-first, invalid symbols stored as (symbol & symbolMask). */
void orc_pack_dense(const orc_config* c, const char* seq, uint64_t num_symbols, uint64_t* out) {
    const uint64_t per_word = 64u / c->symbol_bits;
    const uint64_t symbol_mask = (1ULL << c->symbol_bits) - 1;
    memset(out, 0, orc_dense_word_count(c, num_symbols) * sizeof(uint64_t));
    for (uint64_t i = 0; i < num_symbols; ++i) {
        const uint64_t sym = symbol_at(c, seq, i) & symbol_mask;
        out[i / per_word] |= sym << ((i % per_word) * c->symbol_bits);
    }
}

/* dense_packed.cuh:56-65 */
static uint8_t dense_symbol(const orc_config* c, const uint64_t* packed, uint64_t index) {
    const uint64_t per_word = 64u / c->symbol_bits;
    const uint64_t symbol_mask = (1ULL << c->symbol_bits) - 1;
    return (uint8_t)((packed[index / per_word] >> ((index % per_word) * c->symbol_bits)) &
                     symbol_mask);
}

/* dense_packed.cuh:89-118 then kernels.cuh:106-172 / :40-99 on the decoded tile.
 * A decoded symbol is invalid only if it equals invalidSymbol (never for the
 * built-in alphabets, whose sentinel 0xFF exceeds symbolMask). */
static int dense_window(const orc_config* c, const uint64_t* packed, uint64_t j, uint64_t* out) {
    uint64_t p = 0;
    int valid = 1;
    for (uint32_t i = 0; i < c->k; ++i) {
        const uint8_t sym = dense_symbol(c, packed, j + i);
        if (sym == c->invalid_symbol) valid = 0;
        p = (p << c->symbol_bits) | sym;
    }
    *out = p;
    return valid;
}

void orc_insert_dense(const orc_config* c, const uint64_t* packed, uint64_t num_symbols,
                      uint64_t* words, uint64_t num_shards) {
    const uint64_t n = num_symbols < c->k ? 0 : num_symbols - c->k + 1;
    for (uint64_t j = 0; j < n; ++j) {
        uint64_t p;
        if (!dense_window(c, packed, j, &p)) continue;
        const uint32_t shard = orc_shard_index(orc_window_minimizer_hash(c, p), num_shards);
        uint64_t masks[4];
        orc_window_masks(c, p, masks);
        for (int w = 0; w < 4; ++w) words[(uint64_t)shard * 4u + w] |= masks[w];
    }
}

void orc_contains_dense(const orc_config* c, const uint64_t* packed, uint64_t num_symbols,
                        const uint64_t* words, uint64_t num_shards, uint8_t* out) {
    const uint64_t n = num_symbols < c->k ? 0 : num_symbols - c->k + 1;
    for (uint64_t j = 0; j < n; ++j) {
        uint64_t p;
        if (!dense_window(c, packed, j, &p)) {
            out[j] = 0;
            continue;
        }
        const uint32_t shard = orc_shard_index(orc_window_minimizer_hash(c, p), num_shards);
        out[j] = (uint8_t)contains_packed(c, p, words + (uint64_t)shard * 4u);
    }
}

void orc_count(const orc_config* c, const char* seq, uint64_t bytes, uint64_t num_shards,
               uint64_t* valid_windows, uint64_t* shard_runs) {
    const uint64_t n = orc_kmer_count(c, bytes);
    uint64_t valid = 0, runs = 0;
    int prev_valid = 0;
    uint32_t prev_shard = 0;
    for (uint64_t j = 0; j < n; ++j) {
        if (!orc_window_valid(c, seq, j)) {
            prev_valid = 0;
            continue;
        }
        const uint32_t shard =
            orc_shard_index(orc_window_minimizer_hash(c, orc_pack_window(c, seq, j)), num_shards);
        ++valid;
        if (!prev_valid || shard != prev_shard) ++runs;
        prev_valid = 1;
        prev_shard = shard;
    }
    *valid_windows = valid;
    *shard_runs = runs;
}

/* ------------------------------------------------------------------------- */
/* Fast form: one hash per position, sliding minimum and sliding mask OR.     */
/* ------------------------------------------------------------------------- */

#define ORC_MAX_SPAN 64

typedef struct {
    const orc_config* c;
    const char* seq;
    uint64_t begin, end; /* window range [begin, end) */
    uint64_t* words;
    const uint64_t* cwords;
    uint64_t num_shards;
    uint8_t* out;
    int insert;
} orc_chunk;

static void smer_masks(const orc_config* c, uint64_t smer, uint64_t m[4]) {
    const uint64_t h = orc_hash64(smer);
    m[0] = m[1] = m[2] = m[3] = 0;
    for (uint32_t i = 0; i < c->hash_count; ++i) m[i & 3u] |= 1ULL << orc_bit_address(c, i, h);
}

static void flush_run(uint64_t* words, uint32_t shard, const uint64_t acc[4]) {
    uint64_t* dst = words + (uint64_t)shard * 4u;
    for (int w = 0; w < 4; ++w) {
        if (acc[w]) __atomic_fetch_or(dst + w, acc[w], __ATOMIC_RELAXED);
    }
}

static void* chunk_worker(void* arg) {
    const orc_chunk* ch = (const orc_chunk*)arg;
    const orc_config* c = ch->c;
    if (ch->begin >= ch->end) return NULL;
    const uint32_t bits = c->symbol_bits;
    const uint64_t symbol_mask = (1ULL << bits) - 1;
    const uint64_t mmask = window_mask(c, c->m);
    const uint64_t smask = window_mask(c, c->s);
    const uint32_t w = c->k - c->m + 1;
    const uint32_t f = c->k - c->s + 1;
    const uint64_t shard_mask = ch->num_shards - 1;

    /* monotone deque of (position, hash) over the last w m-mers */
    uint64_t dq_pos[ORC_MAX_SPAN], dq_val[ORC_MAX_SPAN];
    uint32_t dq_head = 0, dq_len = 0;
    /* ring of the last f s-mer masks */
    uint64_t ring[ORC_MAX_SPAN][4];

    uint64_t mm = 0, sm = 0;
    int64_t last_invalid = -1;
    int run_open = 0;
    uint32_t run_shard = 0;
    uint64_t acc[4] = {0, 0, 0, 0};

    const uint64_t first = ch->begin;
    const uint64_t last_symbol = ch->end + c->k - 2;
    for (uint64_t p = first; p <= last_symbol; ++p) {
        const uint8_t sym = symbol_at(c, ch->seq, p);
        if (sym == c->invalid_symbol) last_invalid = (int64_t)p;
        mm = ((mm << bits) | (sym & symbol_mask)) & mmask;
        sm = ((sm << bits) | (sym & symbol_mask)) & smask;
        const uint64_t rel = p - first;
        if (rel + 1 >= c->m) { /* m-mer starting at p-m+1 is complete */
            const uint64_t pos = p - c->m + 1;
            const uint64_t h = orc_minimizer_hash64(mm);
            while (dq_len && dq_val[(dq_head + dq_len - 1) % ORC_MAX_SPAN] > h) --dq_len;
            dq_pos[(dq_head + dq_len) % ORC_MAX_SPAN] = pos;
            dq_val[(dq_head + dq_len) % ORC_MAX_SPAN] = h;
            ++dq_len;
        }
        if (rel + 1 >= c->s) {
            const uint64_t pos = p - c->s + 1;
            smer_masks(c, sm, ring[pos % f]);
        }
        if (rel + 1 < c->k) continue;
        const uint64_t j = p - c->k + 1; /* window complete */
        while (dq_pos[dq_head] < j) {
            dq_head = (dq_head + 1) % ORC_MAX_SPAN;
            --dq_len;
        }
        const int valid = last_invalid < (int64_t)j;
        if (!valid) {
            if (ch->insert) {
                if (run_open) flush_run(ch->words, run_shard, acc);
                run_open = 0;
            } else {
                ch->out[j] = 0;
            }
            continue;
        }
        const uint32_t shard = (uint32_t)(dq_val[dq_head] & shard_mask);
        if (ch->insert) {
            if (run_open && shard == run_shard) {
                const uint64_t* nm = ring[(j + f - 1) % f];
                for (int q = 0; q < 4; ++q) acc[q] |= nm[q];
            } else {
                if (run_open) flush_run(ch->words, run_shard, acc);
                run_open = 1;
                run_shard = shard;
                acc[0] = acc[1] = acc[2] = acc[3] = 0;
                for (uint32_t o = 0; o < f; ++o) {
                    const uint64_t* nm = ring[(j + o) % f];
                    for (int q = 0; q < 4; ++q) acc[q] |= nm[q];
                }
            }
        } else {
            const uint64_t* sw = ch->cwords + (uint64_t)shard * 4u;
            uint64_t miss = 0;
            for (uint32_t o = 0; o < f; ++o) {
                const uint64_t* nm = ring[(j + o) % f];
                for (int q = 0; q < 4; ++q) miss |= nm[q] & ~sw[q];
            }
            ch->out[j] = miss == 0;
        }
        (void)w;
    }
    if (ch->insert && run_open) flush_run(ch->words, run_shard, acc);
    return NULL;
}

int orc_hardware_threads(void) {
    const long n = sysconf(_SC_NPROCESSORS_ONLN);
    return n > 0 ? (int)n : 1;
}

static void run_chunks(const orc_config* c, const char* seq, uint64_t bytes, uint64_t* words,
                       const uint64_t* cwords, uint64_t num_shards, uint8_t* out, int insert,
                       int threads) {
    const uint64_t n = orc_kmer_count(c, bytes);
    if (n == 0) return;
    if (threads <= 0) threads = orc_hardware_threads();
    if ((uint64_t)threads > n) threads = (int)n;
    orc_chunk* chunks = (orc_chunk*)calloc((size_t)threads, sizeof(orc_chunk));
    pthread_t* tids = (pthread_t*)calloc((size_t)threads, sizeof(pthread_t));
    const uint64_t per = (n + (uint64_t)threads - 1) / (uint64_t)threads;
    for (int t = 0; t < threads; ++t) {
        orc_chunk* ch = &chunks[t];
        ch->c = c;
        ch->seq = seq;
        ch->begin = (uint64_t)t * per;
        ch->end = ch->begin + per < n ? ch->begin + per : n;
        ch->words = words;
        ch->cwords = cwords;
        ch->num_shards = num_shards;
        ch->out = out;
        ch->insert = insert;
    }
    for (int t = 1; t < threads; ++t) pthread_create(&tids[t], NULL, chunk_worker, &chunks[t]);
    chunk_worker(&chunks[0]);
    for (int t = 1; t < threads; ++t) pthread_join(tids[t], NULL);
    free(chunks);
    free(tids);
}

void orc_mt_insert(const orc_config* c, const char* seq, uint64_t bytes, uint64_t* words,
                   uint64_t num_shards, int threads) {
    run_chunks(c, seq, bytes, words, words, num_shards, NULL, 1, threads);
}

void orc_mt_contains(const orc_config* c, const char* seq, uint64_t bytes, const uint64_t* words,
                     uint64_t num_shards, uint8_t* out, int threads) {
    run_chunks(c, seq, bytes, NULL, words, num_shards, out, 0, threads);
}
