/*
 * Kuramoto oscillators with a per-oscillator phase lag (libKuramotoLag.so):
 * the Kuramoto plugin's kernels with alpha subtracted inside every sine,
 * bit-exact with src/nemo/cpu/plugins/KuramotoLag.cpp:15-138.
 */
#define NEMO_KURAMOTO_LAG 1
#include "Kuramoto.hip"
