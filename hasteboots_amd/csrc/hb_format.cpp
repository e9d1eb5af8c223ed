// Versioned serialized forms of evaluation keys and LWE ciphertexts (no reference
// counterpart: the reference keeps keys in memory). Layout, little-endian:
//   0  char[4]  magic "HBTF"
//   4  u16      format version (1)
//   6  u16      kind (1 evaluation key, 2 LWE ciphertexts)
//   8  u32 x 8  set, n, N, br_bits, br_len, ks_bits, ks_len, 0
//   40 u64      q
//   48 u64      count (kind 2: ciphertexts; kind 1: 1)
//   56 u64      payload words
//   64 u64      checksum of the payload: h = 0xcbf29ce484222325, per 64-bit word
//               h = (h ^ w) * 0x100000001b3 (FNV-1a over words)
//   72 payload  canonical u64 words: kind 1 the BSK then the KSK in hb_ek_export's layout,
//               kind 2 count * (n + 1) words (a then b per ciphertext)
// Reading checks the magic, version, kind, that the parameters equal the set's, the
// payload length, the checksum and that every word is below q.
#include <cstring>
#include <vector>

#include "hb_internal.hpp"

using namespace hb;

namespace {

constexpr uint16_t kVersion = 1;
constexpr size_t kHeader = 72;

uint64_t words_hash(const uint64_t* w, size_t n) {
  uint64_t h = 0xcbf29ce484222325ull;
  for (size_t k = 0; k < n; k++) h = (h ^ w[k]) * 0x100000001b3ull;
  return h;
}

struct Header {
  uint16_t kind;
  ParamSet p;
  uint64_t count, payload_words, checksum;
};

void put_header(uint8_t* out, uint16_t kind, const ParamSet& p, uint64_t count,
                uint64_t payload_words, uint64_t checksum) {
  std::memset(out, 0, kHeader);
  std::memcpy(out, "HBTF", 4);
  std::memcpy(out + 4, &kVersion, 2);
  std::memcpy(out + 6, &kind, 2);
  const uint32_t f[8] = {(uint32_t)p.set, p.n, p.N, p.br_bits, p.br_len, p.ks_bits, p.ks_len, 0};
  std::memcpy(out + 8, f, 32);
  std::memcpy(out + 40, &p.q, 8);
  std::memcpy(out + 48, &count, 8);
  std::memcpy(out + 56, &payload_words, 8);
  std::memcpy(out + 64, &checksum, 8);
}

int get_header(const uint8_t* in, size_t size, uint16_t want_kind, Header* h) {
  if (!in || size < kHeader) return fail(HB_ERR_ARG, "serialized data shorter than its header");
  if (std::memcmp(in, "HBTF", 4) != 0) return fail(HB_ERR_ARG, "not a serialized hb object");
  uint16_t version, kind;
  std::memcpy(&version, in + 4, 2);
  std::memcpy(&kind, in + 6, 2);
  if (version != kVersion) return fail(HB_ERR_ARG, "format version %u not supported", version);
  if (kind != want_kind) return fail(HB_ERR_ARG, "serialized kind %u, expected %u", kind, want_kind);
  uint32_t f[8];
  std::memcpy(f, in + 8, 32);
  if (!param_set((int)f[0], &h->p)) return fail(HB_ERR_ARG, "unknown parameter set %u", f[0]);
  const ParamSet& p = h->p;
  uint64_t q;
  std::memcpy(&q, in + 40, 8);
  if (f[1] != p.n || f[2] != p.N || f[3] != p.br_bits || f[4] != p.br_len || f[5] != p.ks_bits ||
      f[6] != p.ks_len || q != p.q)
    return fail(HB_ERR_ARG, "serialized parameters differ from set %u's", f[0]);
  std::memcpy(&h->count, in + 48, 8);
  std::memcpy(&h->payload_words, in + 56, 8);
  std::memcpy(&h->checksum, in + 64, 8);
  if (h->payload_words > (size - kHeader) / 8 || size != kHeader + h->payload_words * 8)
    return fail(HB_ERR_ARG, "serialized length does not match the header");
  h->kind = kind;
  return HB_OK;
}

int check_payload(const Header& h, const uint64_t* w) {
  if (words_hash(w, h.payload_words) != h.checksum) return fail(HB_ERR_ARG, "checksum mismatch");
  for (uint64_t k = 0; k < h.payload_words; k++)
    if (w[k] >= h.p.q) return fail(HB_ERR_ARG, "word %llu is not canonical", (unsigned long long)k);
  return HB_OK;
}

}  // namespace

extern "C" {

int hb_keys_serialize(int set, const uint64_t* bsk, const uint64_t* ksk, void* buf, size_t cap,
                      size_t* size) {
  ParamSet p;
  if (!param_set(set, &p)) return fail(HB_ERR_ARG, "unknown parameter set %d", set);
  if (!bsk || !ksk || !size) return fail(HB_ERR_ARG, "null argument");
  const uint64_t nb = bsk_words(p), nk = ksk_words(p);
  *size = kHeader + (nb + nk) * 8;
  if (!buf) return HB_OK;
  if (cap < *size) return fail(HB_ERR_CAPACITY, "buffer of %zu bytes, %zu needed", cap, *size);
  for (uint64_t k = 0; k < nb; k++)
    if (bsk[k] >= p.q) return fail(HB_ERR_ARG, "bsk word %llu is not canonical", (unsigned long long)k);
  for (uint64_t k = 0; k < nk; k++)
    if (ksk[k] >= p.q) return fail(HB_ERR_ARG, "ksk word %llu is not canonical", (unsigned long long)k);
  uint8_t* o = static_cast<uint8_t*>(buf);
  std::memcpy(o + kHeader, bsk, nb * 8);
  std::memcpy(o + kHeader + nb * 8, ksk, nk * 8);
  uint64_t h = 0xcbf29ce484222325ull;
  for (uint64_t k = 0; k < nb; k++) h = (h ^ bsk[k]) * 0x100000001b3ull;
  for (uint64_t k = 0; k < nk; k++) h = (h ^ ksk[k]) * 0x100000001b3ull;
  put_header(o, 1, p, 1, nb + nk, h);
  return HB_OK;
}

int hb_keys_deserialize(const void* buf, size_t size, int* set, uint64_t* bsk, uint64_t* ksk) {
  Header h;
  int rc = get_header(static_cast<const uint8_t*>(buf), size, 1, &h);
  if (rc) return rc;
  const uint64_t nb = bsk_words(h.p), nk = ksk_words(h.p);
  if (h.count != 1 || h.payload_words != nb + nk)
    return fail(HB_ERR_ARG, "key payload of %llu words, %llu expected",
                (unsigned long long)h.payload_words, (unsigned long long)(nb + nk));
  std::vector<uint64_t> w(h.payload_words);
  std::memcpy(w.data(), static_cast<const uint8_t*>(buf) + kHeader, w.size() * 8);
  rc = check_payload(h, w.data());
  if (rc) return rc;
  if (set) *set = h.p.set;
  if (bsk) std::memcpy(bsk, w.data(), nb * 8);
  if (ksk) std::memcpy(ksk, w.data() + nb, nk * 8);
  return HB_OK;
}

int hb_ek_serialize(const hb_ek_t* ek, void* buf, size_t cap, size_t* size) {
  if (!ek) return fail(HB_ERR_ARG, "null ek");
  return hb_keys_serialize(ek->p.set, ek->bsk_host.data(), ek->ksk_host.data(), buf, cap, size);
}

int hb_ek_deserialize(const void* buf, size_t size, int device, hb_ek_t** out) {
  if (!out) return fail(HB_ERR_ARG, "null out");
  Header h;
  int rc = get_header(static_cast<const uint8_t*>(buf), size, 1, &h);
  if (rc) return rc;
  std::vector<uint64_t> bsk(bsk_words(h.p)), ksk(ksk_words(h.p));
  int set = 0;
  rc = hb_keys_deserialize(buf, size, &set, bsk.data(), ksk.data());
  if (rc) return rc;
  return hb_ek_import(set, bsk.data(), ksk.data(), device, out);
}

int hb_lwe_serialize(int set, size_t count, const uint64_t* lwes, void* buf, size_t cap,
                     size_t* size) {
  ParamSet p;
  if (!param_set(set, &p)) return fail(HB_ERR_ARG, "unknown parameter set %d", set);
  if ((!lwes && count) || !size) return fail(HB_ERR_ARG, "null argument");
  const uint64_t words = (uint64_t)count * (p.n + 1);
  *size = kHeader + words * 8;
  if (!buf) return HB_OK;
  if (cap < *size) return fail(HB_ERR_CAPACITY, "buffer of %zu bytes, %zu needed", cap, *size);
  for (uint64_t k = 0; k < words; k++)
    if (lwes[k] >= p.q) return fail(HB_ERR_ARG, "word %llu is not canonical", (unsigned long long)k);
  uint8_t* o = static_cast<uint8_t*>(buf);
  if (words) std::memcpy(o + kHeader, lwes, words * 8);
  put_header(o, 2, p, count, words, words_hash(lwes, words));
  return HB_OK;
}

int hb_lwe_deserialize(const void* buf, size_t size, int* set, size_t* count, uint64_t* lwes,
                       size_t cap_words) {
  Header h;
  int rc = get_header(static_cast<const uint8_t*>(buf), size, 2, &h);
  if (rc) return rc;
  if (h.payload_words != h.count * (h.p.n + 1))
    return fail(HB_ERR_ARG, "ciphertext payload does not match its count");
  if (set) *set = h.p.set;
  if (count) *count = (size_t)h.count;
  if (!lwes) return HB_OK;
  if (cap_words < h.payload_words)
    return fail(HB_ERR_CAPACITY, "room for %zu words, %llu needed", cap_words,
                (unsigned long long)h.payload_words);
  std::vector<uint64_t> w(h.payload_words);
  if (!w.empty()) std::memcpy(w.data(), static_cast<const uint8_t*>(buf) + kHeader, w.size() * 8);
  rc = check_payload(h, w.data());
  if (rc) return rc;
  if (!w.empty()) std::memcpy(lwes, w.data(), w.size() * 8);
  return HB_OK;
}

}  // extern "C"
