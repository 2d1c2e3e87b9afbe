// oracle_rt_abi.cpp — C ABI over the oracle runtime (oracle/svm_oracle_rt.h) so
// tests can check Hash / StrHash / StrCmp / BuildHashTable / Probe directly.
// TEST INFRASTRUCTURE ONLY.
#include "svm_oracle_rt.h"

extern "C" {

unsigned ort_hash(const char* p, int n) { return Hash(p, n); }
unsigned ort_strhash(const char* p, int n) { return StrHash(p, n); }
int ort_strcmp(const char* a, const char* b, int n) { return StrCmp(a, b, n); }

// Builds a table over int keys (row order) and returns the dropped-insert count.
uint64 ort_build_int(const int* keys, unsigned n, int* ht, unsigned ht_size) {
  return BuildHashTable<int>(keys, n, 4, ht, ht_size);
}
uint64 ort_build_str(const char* keys, unsigned n, unsigned width, int* ht, unsigned ht_size) {
  return BuildHashTable<char>(keys, n, width, ht, ht_size);
}
int ort_probe_int(int key, const int* tab, const int* ht, unsigned ht_size) {
  return Probe<int>(&key, 4, tab, ht, ht_size);
}
int ort_probe_str(const char* key, unsigned width, const char* tab, const int* ht, unsigned ht_size) {
  return Probe<char>(key, width, tab, ht, ht_size);
}

}  // extern "C"
