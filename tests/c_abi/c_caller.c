/* A plain-C caller of include/bbwadg.h: checks that the header compiles as C and that the
 * library's argument checks return the documented error classes before any device work
 * (run by tests/test_c_caller.py; needs no GPU). */
#include <stdio.h>
#include <string.h>

#include "bbwadg.h"

static int failures = 0;

static void expect(int got, int want, const char *what) {
  if (got != want) {
    fprintf(stderr, "%s: got %d, want %d\n", what, got, want);
    ++failures;
  }
}

int main(void) {
  bbw_ctx *ctx = (bbw_ctx *)0x1;
  bbw_config cfg;
  memset(&cfg, 0, sizeof cfg);

  cfg.N = 10; cfg.Nfields = 4; cfg.K = 10; cfg.Nq = 1331;
  expect(bbw_create(&cfg, &ctx), BBW_ERR_UNSUPPORTED, "order 10");
  expect(ctx == NULL, 1, "context cleared on failure");
  cfg.N = 5; cfg.Nfields = 5; cfg.Nq = 95;
  expect(bbw_create(&cfg, &ctx), BBW_ERR_UNSUPPORTED, "5 fields");
  cfg.N = 7; cfg.Nfields = 9; cfg.Nq = 214;
  expect(bbw_create(&cfg, &ctx), BBW_ERR_UNSUPPORTED, "elastic at N = 7");
  cfg.N = 5; cfg.Nfields = 4; cfg.Nq = 94;
  expect(bbw_create(&cfg, &ctx), BBW_ERR_ARG, "wrong Nq");
  cfg.Nq = 95; cfg.K = 0;
  expect(bbw_create(&cfg, &ctx), BBW_ERR_ARG, "K = 0");
  cfg.K = 1 << 30;
  expect(bbw_create(&cfg, &ctx), BBW_ERR_ARG, "state larger than 2^31 floats");
  expect(bbw_create(NULL, &ctx), BBW_ERR_ARG, "null config");
  expect(bbw_create(&cfg, NULL), BBW_ERR_ARG, "null output");

  /* null contexts */
  expect(bbw_step(NULL, 1, 1e-3, 0.0, BBW_MODE_PLAIN, 0), BBW_ERR_ARG, "step(NULL)");
  expect(bbw_upload_slice_lift(NULL, NULL, NULL, NULL, NULL), BBW_ERR_ARG, "slice lift(NULL)");
  expect(bbw_set_option(NULL, BBW_OPT_PLAIN_IMPL, BBW_PLAIN_SLICE), BBW_ERR_ARG, "option(NULL)");
  expect(strcmp(bbw_last_error(NULL), "null context") == 0, 1, "last_error(NULL)");
  bbw_destroy(NULL);

  bbw_batch *b = (bbw_batch *)0x1;
  expect(bbw_batch_create(7, 0, &b), BBW_ERR_UNSUPPORTED, "batch at N = 7");

  if (failures == 0) printf("c caller: all checks passed\n");
  return failures == 0 ? 0 : 1;
}
