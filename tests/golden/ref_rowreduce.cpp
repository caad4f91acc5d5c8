// Test-only harness for gen_ref_vectors.py: calls the reference runtime's
// corrWith and aggregrate (KunQuant cpp/Kun/RunGraph.hpp:15-20, 55-75), which
// the reference's C API does not export, with plain pointers from ctypes.
// Built against the reference's headers and its runtime in oracle/_ref.
#include <Kun/Context.hpp>
#include <Kun/RunGraph.hpp>

#include <vector>

extern "C" int ref_corr_with(int layout_sts, int rank_inputs, int n, float **inputs, float *target, float **outs,
                             size_t num_stocks, size_t total_time, size_t cur_time, size_t length) {
    auto ex = kun::createSingleThreadExecutor();
    std::vector<float *> in(inputs, inputs + n), out(outs, outs + n);
    kun::corrWith(ex, layout_sts ? kun::MemoryLayout::STs : kun::MemoryLayout::TS, rank_inputs != 0, in, target, out,
                  num_stocks, total_time, cur_time, length);
    return 0;
}

// outs: n x 7 pointers (sum, min, max, first, last, count, mean; null = not wanted)
extern "C" int ref_aggregrate(int n, float **inputs, float **labels, int is_double, float **outs, size_t num_stocks,
                              size_t total_time, size_t cur_time, size_t length) {
    auto ex = kun::createSingleThreadExecutor();
    std::vector<kun::AggregrationOutput> o(n);
    for (int i = 0; i < n; i++)
        for (int j = 0; j < kun::AGGREGRATION_NUM_KINDS; j++) o[i].buffers[j] = outs[i * kun::AGGREGRATION_NUM_KINDS + j];
    kun::aggregrate(ex, n, inputs, labels, is_double ? kun::Datatype::Double : kun::Datatype::Float, o.data(),
                    num_stocks, total_time, cur_time, length);
    return 0;
}
