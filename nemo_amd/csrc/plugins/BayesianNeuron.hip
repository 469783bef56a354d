/*
 * BayesianNeuron plugin (libBayesianNeuron.so), bit-exact with src/nemo/cpu/plugins/BayesianNeuron.cpp:14-93
 * (the update is in variants.h).
 */
#define NEMO_VARIANT VARIANT_BAYESIAN
#include "variants.h"
