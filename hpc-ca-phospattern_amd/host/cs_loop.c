/*
 * Time-loop bookkeeping of the reference driver, kept on the host so the
 * device can run long step ranges without synchronising every step.
 */
#include <math.h>
#include <stdint.h>
#include "cs_host.h"

uint64_t cs_save_interval(double dt_save, double dt)
{
  /* save_interval = max_int(1, floor(dt_save/dt)) (calcium_sparks.c:168);
   * max_int takes ints, so the floor is truncated to int first. */
  int v = (int)floor(dt_save / dt);
  return (uint64_t)(v > 1 ? v : 1);
}

uint64_t cs_count_steps(double tstop, double dt)
{
  /* for (t=0., time_ind=1; t<tstop; t+=dt, time_ind++)  (:190) */
  uint64_t n = 0;
  for (double t = 0.0; t < tstop; t += dt) n++;
  return n;
}

void cs_abort_init(cs_abort_state* a, double tabort, int have_discrete)
{
  a->tabort = tabort;
  a->last_closed = -1.0;
  a->all_closed = 0;
  a->active = (tabort != 0.0) && have_discrete;
}

void cs_abort_snapshot(cs_abort_state* a, int all_closed)
{
  a->all_closed = all_closed;
}

int cs_abort_check(cs_abort_state* a, double t)
{
  /* species.c:3396-3487: inactive unless tabort != 0 and there are
   * discrete boundaries; "not all closed" resets the clock; a negative
   * tabort never aborts. */
  if (!a->active) return 0;
  if (!a->all_closed) {
    a->last_closed = -1.0;
    return 0;
  }
  if (a->tabort >= 0) {
    if (a->last_closed < 0) a->last_closed = t;
    if (t - a->last_closed >= a->tabort) return 1;
  }
  return 0;
}
