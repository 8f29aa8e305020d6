// trajsim_compat.hpp -- the reference's 12 entry points, re-exported with
// their original C++ signatures so a harness written against
// 1li2yang3/CudaRuntime1 (main.cu) relinks against libtrajsim.so unchanged.
//
// Semantics follow the reference: t1 holds num_t queries of length n, t2 the
// database (num_t, or num_t2, trajectories of length m), h_results[i] the best
// value for query i; time out-params and return values are milliseconds. All
// of them run on the GPU through the C ABI in trajsim.h (the "_cpu" names are
// kept for source compatibility: their results equal the reference's CPU
// path). Errors print a message and exit(1), like the reference's CHECK
// (common.cuh:14-22).
#ifndef TRAJSIM_COMPAT_HPP
#define TRAJSIM_COMPAT_HPP

// common.cuh:9-11
struct Point {
  float x, y;
};

// dtw.cuh:6-7
void launch_dtw_batch_gpu_knn(const Point* h_t1, const Point* h_t2, float* h_results, int num_t, int n, int m,
                              float& gpu_time);
float launch_dtw_batch_cpu(const Point* h_t1, const Point* h_t2, float* h_results, int num_t, int n, int m);
// frechet.cuh:6,8
void launch_frechet_batch_gpu_wavefront(const Point* h_t1, const Point* h_t2, float* h_results, int num_t, int n,
                                        int m, float& gpu_time);
float launch_frechet_batch_cpu(const Point* h_t1, const Point* h_t2, float* h_results, int num_t, int n, int m);
// lcss.cuh:6-8
void launch_lcss_batch_gpu_wavefront_knn(const Point* h_t1, const Point* h_t2, float* h_results, int num_t, int n,
                                         int m, float epsilon, float& gpu_time);
float launch_lcss_batch_cpu(const Point* h_t1, const Point* h_t2, float* h_results, int num_t, int n, int m,
                            float epsilon);
// hausdorff.cuh:7-19
float launch_hausdorff_batch_cpu_rtree(const Point* h_t1, int num_t1, int n, const Point* h_t2, int num_t2, int m,
                                       float* h_results, int top_k = 10);
float run_hausdorff_rtree_gpu_pipeline(const Point* h_t1, int num_t1, int n, const Point* h_t2, int num_t2, int m,
                                       float* h_results, int top_k = 10);
void launch_hausdorff_rtree_gpu(const Point* h_t1, int num_t1, int n, const Point* h_t2, int num_t2, int m,
                                const int* h_candidate_indices, int top_k, float* h_results, float& gpu_time);
// euclidean.cuh:7-13
float launch_euclidean_batch_cpu_rtree(const Point* h_t1, const Point* h_t2, float* h_results, int num_t, int n);
void launch_euclidean_batch_gpu_rtree_exact(const Point* h_t1, const Point* h_t2, float* h_results, int num_t, int n,
                                            float& gpu_time);
void execute_rtree_csr_kernel_on_gpu(const Point* h_t1, const Point* h_t2, const int* h_candidates,
                                     const int* h_offsets, float* h_results, int num_t, int n,
                                     int total_candidates, float& gpu_time);

#endif
