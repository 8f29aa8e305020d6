// ref_shim.cpp -- C-linkage driver over the reference's own CPU functions.
//
// TEST INFRASTRUCTURE ONLY: linked with the reference's dtw.cpp, frechet.cpp
// and lcss.cpp (compiled from /root/reference by oracle/Makefile) into
// oracle/_ref/libref.so, so that the oracle can be checked against the
// reference itself. Nothing in the product links this.
//
// The prototypes below restate the reference's declarations
// (dtw.cuh:7, frechet.cuh:8, lcss.cuh:8) and its point type (common.cuh:9-11)
// so this file needs none of the reference headers.
struct Point { float x, y; };

float launch_dtw_batch_cpu(const Point* h_t1, const Point* h_t2, float* h_results, int num_t, int n, int m);
float launch_frechet_batch_cpu(const Point* h_t1, const Point* h_t2, float* h_results, int num_t, int n, int m);
float launch_lcss_batch_cpu(const Point* h_t1, const Point* h_t2, float* h_results, int num_t, int n, int m, float epsilon);

extern "C" {
float ref_dtw_cpu(const float* t1, const float* t2, float* out, int num_t, int n, int m) {
  return launch_dtw_batch_cpu((const Point*)t1, (const Point*)t2, out, num_t, n, m);
}
float ref_frechet_cpu(const float* t1, const float* t2, float* out, int num_t, int n, int m) {
  return launch_frechet_batch_cpu((const Point*)t1, (const Point*)t2, out, num_t, n, m);
}
float ref_lcss_cpu(const float* t1, const float* t2, float* out, int num_t, int n, int m, float eps) {
  return launch_lcss_batch_cpu((const Point*)t1, (const Point*)t2, out, num_t, n, m, eps);
}
}
