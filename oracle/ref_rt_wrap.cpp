// ref_rt_wrap.cpp — C ABI over the reference's own CPU runtime helpers, compiled
// from the reference sources where they lie (com/cpp/define.hpp and
// com/cpp/cpu-runtime.hpp) into oracle/_ref/libref_rt.so by oracle/Makefile.
// TEST INFRASTRUCTURE ONLY: tests use it to check the oracle's restatement of
// Hash / StrHash / StrCmp / InsertHashTable / Probe on many inputs.
#include <cstdint>
#include <cstdlib>

#include "define.hpp"
#include "cpu-runtime.hpp"

extern "C" {

unsigned ref_hash(const char* p, int n) { return Hash(p, n); }
unsigned ref_strhash(const char* p, int n) { return StrHash(p, n); }
int ref_strcmp(const char* a, const char* b, int n) { return StrCmp(a, b, n); }

// The fill + row-order insert loop of BuildHashTable (com/cpp/com.hpp:168-174),
// counting the inserts InsertHashTable reports as failed.
unsigned long long ref_build_int(int* keys, unsigned n, int* ht, unsigned ht_size) {
  for (unsigned s = 0; s < ht_size; ++s) ht[s] = -1;
  unsigned long long dropped = 0;
  for (unsigned r = 0; r < n; ++r)
    if (!InsertHashTable<int>(keys, (int)r, 4, ht, ht_size)) ++dropped;
  return dropped;
}
unsigned long long ref_build_str(char* keys, unsigned n, unsigned width, int* ht, unsigned ht_size) {
  for (unsigned s = 0; s < ht_size; ++s) ht[s] = -1;
  unsigned long long dropped = 0;
  for (unsigned r = 0; r < n; ++r)
    if (!InsertHashTable<char>(keys, (int)r, width, ht, ht_size)) ++dropped;
  return dropped;
}
int ref_probe_int(int key, int* tab, int* ht, unsigned ht_size) { return Probe<int>(&key, 4, tab, ht, ht_size); }
int ref_probe_str(char* key, unsigned width, char* tab, int* ht, unsigned ht_size) {
  return Probe<char>(key, width, tab, ht, ht_size);
}

}  // extern "C"
