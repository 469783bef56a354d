/*
 * InhomogeneousPoissonSource plugin (libInhomogeneousPoissonSource.so), bit-exact with src/nemo/cpu/plugins/InhomogeneousPoissonSource.cpp:6-27
 * (the update is in variants.h).
 */
#define NEMO_VARIANT VARIANT_INHOMOGENEOUS_POISSON
#include "variants.h"
