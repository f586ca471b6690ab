/*
 * orc_internal.h -- iteration bookkeeping shared by the oracle's solver
 * restatements (solver/iterative.hpp:478-507, 748-885).  Test infrastructure.
 */
#ifndef ORC_INTERNAL_H
#define ORC_INTERNAL_H
#include "feat_oracle.h"

#define ORC_HIDDEN __attribute__((visibility("hidden")))

typedef struct {
  const orc_iter_cfg_t* c;
  double def_init, def_cur, def_prev;
  uint64_t num_iter, num_stag_iter;
  double* hist;
} itstate_t;

ORC_HIDDEN int set_initial_defect(itstate_t* s, double def);
ORC_HIDDEN int analyse_defect(itstate_t* s, uint64_t it, double cur, double prev, int check_stag);
ORC_HIDDEN int set_new_defect(itstate_t* s, const double* def, uint64_t n);
ORC_HIDDEN int status_success(int st);
#endif
