// server.cpp -- continuous-batching request loop (placeholder until the
// serving scheduler lands).
#include "../../include/trajsim.h"

extern "C" {
void ts_server_config_default(ts_server_config* c) {
  if (!c) return;
  c->streams = 4;
  c->max_batch = 256;
  c->max_queue = 65536;
  c->mem_reserve = 0;
  c->fake_free_mem = 0;
  c->batch_wait_us = 200;
}
ts_status ts_server_create(const ts_set*, const ts_server_config*, ts_server**) { return TS_ERR_UNSUPPORTED; }
ts_status ts_server_submit(ts_server*, const ts_point*, int32_t, const ts_params*, uint64_t*) { return TS_ERR_UNSUPPORTED; }
ts_status ts_server_wait(ts_server*, uint64_t, int32_t, int32_t*, float*, double*) { return TS_ERR_UNSUPPORTED; }
ts_status ts_server_cancel(ts_server*, uint64_t) { return TS_ERR_UNSUPPORTED; }
ts_status ts_server_counters(ts_server*, int64_t*) { return TS_ERR_UNSUPPORTED; }
ts_status ts_server_destroy(ts_server*) { return TS_OK; }
}
