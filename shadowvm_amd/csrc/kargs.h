// kargs.h — launch-argument block shared by the host and every generated kernel.
// The same text is compiled into libsvm and handed to hipRTC, so both sides agree
// on the layout by construction.
#pragma once

// Slot map of ptr[]:
//   0..23  source (fact) columns by column index
//   24..31 packed join payload maps, by join position
//   32..47 dimension columns read by expressions, assigned by the generator
//   48..50 per-dimension-row key code columns (group keys read through a row id)
//   51     the narrow copy of the tier-0 key column (KernelShape::packed)
//   54..63 projected output columns of a select pipeline
// out[]: 0 group sums (int64), 1 group flags (bit per group), 2 scalar sums,
//        6 select tile offsets (u64), 7 select tile counts (u32)
// iv[]: 4j+0 kmin, 4j+1 key range, 4j+2 map words, 4j+3 LDS word offset of join j;
//       32+4q+0 fact-key minimum, 32+4q+1 stride, 32+4q+2 cardinality of key q;
//       44 group count, 45 LDS word offset of the aggregation table;
//       48 the minimum added back to the values of the narrow copy
#define SVM_KARGS_DEF                                                                     \
  struct SvmArgs {                                                                        \
    const void* ptr[64];                                                                  \
    void* out[8];                                                                         \
    long long row0;                                                                       \
    long long rows;                                                                       \
    int iv[50];                                                                           \
    int pi[64];                                                                           \
    const char* sblob;                                                                    \
  };

SVM_KARGS_DEF

// What one host task of the CPU engine writes (the generated host code and libsvm
// share this text, as they share SvmArgs):
//   acc        scalar sums (K = 0)
//   gsum/gflag dense group table (G x M int64) and its flags; `shared`: the table is
//              updated by concurrent tasks (atomic adds), else private to the task
//   ekeys/evals/nentries  packed group keys and metrics of a wide group space, one
//              entry per passing row
//   sel/nsel   select: the task's projected columns and its row count
#define SVM_HOSTOUT_DEF                                                                   \
  struct SvmHostOut {                                                                     \
    long long acc[8];                                                                     \
    unsigned long long* gsum;                                                             \
    unsigned* gflag;                                                                      \
    int shared;                                                                           \
    unsigned long long* ekeys;                                                            \
    long long* evals;                                                                     \
    unsigned long long nentries;                                                          \
    char* sel[10];                                                                        \
    unsigned long long nsel;                                                              \
  };

SVM_HOSTOUT_DEF

#define SVM_KSTR2(x) #x
#define SVM_KSTR(x) SVM_KSTR2(x)
#define SVM_KARGS_TEXT SVM_KSTR(SVM_KARGS_DEF)
#define SVM_HOSTOUT_TEXT SVM_KSTR(SVM_HOSTOUT_DEF)

#define SVM_SLOT_FACT 0
#define SVM_SLOT_MAP 24
#define SVM_SLOT_DCOL 32
#define SVM_SLOT_CODE 48
#define SVM_SLOT_PACK 51
#define SVM_MAX_PACKED 1
#define SVM_SLOT_OUT 54
#define SVM_MAX_SELECT 10
#define SVM_MAX_JOIN 8
#define SVM_MAX_LIT 64
