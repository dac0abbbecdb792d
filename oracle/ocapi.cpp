// ocapi.cpp — ORACLE C ABI (test infrastructure only): the product's C-ABI
// glue instantiated over the CPU restatement, exported as oracle_* so tests
// can run the same calls through both libraries and compare the results.
#include "otree.hpp"

#define CUACD_API(name) oracle_##name
#define CUACD_MAKE_BACKEND(cfg) (new orc::OracleBackend(cfg))
#include "../cuacd-web_amd/csrc/host/capi_body.inc"
