// Operator factories with the reference's names and call shapes
// (fused_factory.hpp:34-86, labeling_factory.hpp:33-85,
// mass_update_factory.hpp:36-94, centroid_update_factory.hpp:37-125) and the
// helper operators MatrixBinaryOp (matrix_binary_op.hpp:30-219) and
// ReduceVectorParcol (reduce_vector_parcol.hpp:30-167).
//
// Each factory maps its strategy string to a C-ABI strategy id (unknown
// strings throw std::invalid_argument, as the reference factories do) and
// returns a std::function that enqueues the operator on a Queue after the
// wait list, records a timing event pair into the DataPoint and returns a
// completion Event.  Centroid sums are accumulated in an FP64 buffer
// (AccT = double) rather than PointT.
#pragma once

#include <functional>
#include <string>
#include <type_traits>

#include "configuration.hpp"
#include "device.hpp"
#include "measurement.hpp"

namespace Clustering {

using AccT = double;

template <typename PointT, typename LabelT, typename MassT>
constexpr void require_supported_types()
{
    static_assert((std::is_same<PointT, float>::value && std::is_same<LabelT, uint32_t>::value &&
                   std::is_same<MassT, uint32_t>::value) ||
                      (std::is_same<PointT, double>::value && std::is_same<LabelT, uint64_t>::value &&
                       std::is_same<MassT, uint64_t>::value),
                  "the gfx950 kernels implement the float/uint32/uint32 and double/uint64/uint64 "
                  "type triples (bench.cpp:469-503)");
}

// The point type that goes with a label type (uint32 -> float, uint64 -> double).
template <typename LabelT>
using PointOfLabel = typename std::conditional<std::is_same<LabelT, uint64_t>::value, double, float>::type;

// The C-ABI entry point of each operator for either type triple.
namespace abi {
inline int lloyd_step(clk_context *c, const clk_kernel_config *k, size_t d, size_t n, size_t kk, const float *p,
                      const float *oc, float *nc, double *s, uint32_t *l, uint32_t *m, double *sh, hipStream_t st)
{
    return clk_lloyd_step(c, k, d, n, kk, p, oc, nc, s, l, m, sh, st);
}
inline int lloyd_step(clk_context *c, const clk_kernel_config *k, size_t d, size_t n, size_t kk, const double *p,
                      const double *oc, double *nc, double *s, uint64_t *l, uint64_t *m, double *sh, hipStream_t st)
{
    return clk_lloyd_step_f64(c, k, d, n, kk, p, oc, nc, s, l, m, sh, st);
}
inline int fused(clk_context *c, const clk_kernel_config *k, size_t d, size_t n, size_t kk, const float *p,
                 const float *oc, double *s, uint32_t *l, uint32_t *m, hipStream_t st)
{
    return clk_fused(c, k, d, n, kk, p, oc, s, l, m, st);
}
inline int fused(clk_context *c, const clk_kernel_config *k, size_t d, size_t n, size_t kk, const double *p,
                 const double *oc, double *s, uint64_t *l, uint64_t *m, hipStream_t st)
{
    return clk_fused_f64(c, k, d, n, kk, p, oc, s, l, m, st);
}
inline int labeling(clk_context *c, const clk_kernel_config *k, size_t d, size_t n, size_t kk, const float *p,
                    const float *cen, uint32_t *l, hipStream_t st)
{
    return clk_labeling(c, k, d, n, kk, p, cen, l, st);
}
inline int labeling(clk_context *c, const clk_kernel_config *k, size_t d, size_t n, size_t kk, const double *p,
                    const double *cen, uint64_t *l, hipStream_t st)
{
    return clk_labeling_f64(c, k, d, n, kk, p, cen, l, st);
}
inline int mass_update(clk_context *c, const clk_kernel_config *k, size_t n, size_t kk, const uint32_t *l,
                       uint32_t *m, hipStream_t st)
{
    return clk_mass_update(c, k, n, kk, l, m, st);
}
inline int mass_update(clk_context *c, const clk_kernel_config *k, size_t n, size_t kk, const uint64_t *l,
                       uint64_t *m, hipStream_t st)
{
    return clk_mass_update_u64(c, k, n, kk, l, m, st);
}
inline int centroid_update(clk_context *c, const clk_kernel_config *k, size_t d, size_t n, size_t kk,
                           const float *p, double *s, const uint32_t *l, const uint32_t *m, hipStream_t st)
{
    return clk_centroid_update(c, k, d, n, kk, p, s, l, m, st);
}
inline int centroid_update(clk_context *c, const clk_kernel_config *k, size_t d, size_t n, size_t kk,
                           const double *p, double *s, const uint64_t *l, const uint64_t *m, hipStream_t st)
{
    return clk_centroid_update_f64(c, k, d, n, kk, p, s, l, m, st);
}
inline int divide(clk_context *c, size_t d, size_t k, const double *s, const uint32_t *m, float *cen,
                  const float *old, double *sh, hipStream_t st)
{
    return clk_matrix_divide_row(c, d, k, s, m, cen, old, sh, st);
}
inline int divide(clk_context *c, size_t d, size_t k, const double *s, const uint64_t *m, double *cen,
                  const double *old, double *sh, hipStream_t st)
{
    return clk_matrix_divide_f64(c, d, k, s, m, cen, old, sh, st);
}
inline int matrix_add(clk_context *c, size_t count, uint32_t *a, const uint32_t *b, hipStream_t st)
{
    return clk_matrix_add_u32(c, count, a, b, st);
}
inline int matrix_add(clk_context *c, size_t count, uint64_t *a, const uint64_t *b, hipStream_t st)
{
    return clk_matrix_add_u64(c, count, a, b, st);
}
inline int matrix_add(clk_context *c, size_t count, double *a, const double *b, hipStream_t st)
{
    return clk_matrix_add_f64(c, count, a, b, st);
}
}  // namespace abi

// DataPoint name of a fused strategy id (the reference's kernel class names).
inline std::string fused_kernel_name(int id)
{
    return id == CLK_FUSED_FEATURE_SUM           ? "FusedFeatureSum"
           : id == CLK_FUSED_CLUSTER_MERGE_BOUNDED ? "FusedClusterMergeBounded"
                                                   : "FusedClusterMerge";
}

inline void fill_sizes(clk_kernel_config &k, const size_t g[3], const size_t l[3])
{
    for (int i = 0; i < 3; ++i) {
        k.global_size[i] = g[i];
        k.local_size[i] = l[i];
    }
}

// Enqueue `fn(stream)` after `events`, timed into `dp`; returns completion.
template <typename F>
Event timed_enqueue(const Queue &q, Measurement::DataPoint &dp, const WaitList &events, F &&fn)
{
    enqueue_wait(q.get(), events);
    Measurement::TimedEvent &te = dp.add_event();
    te = Measurement::TimedEvent::begin(q.get());
    fn(q.get());
    te.finish(q.get());
    return Event::record(q.get());
}

template <typename T>
void require_len(const BufferIterator<T> &b, const BufferIterator<T> &e, size_t n, const char *what)
{
    if ((size_t)(e - b) != n || b.get_index() != 0)
        throw std::length_error(std::string(what) + ": buffer size/offset mismatch");
}

// ---------------------------------------------------------------- fused --

template <typename PointT, typename LabelT, typename MassT, bool ColMajor = true>
class FusedFactory {
public:
    using FusedFunction = std::function<Event(
        const Queue &queue, size_t num_features, size_t num_points, size_t num_clusters,
        BufferIterator<PointT> points_begin, BufferIterator<PointT> points_end,
        BufferIterator<PointT> old_centroids_begin, BufferIterator<PointT> old_centroids_end,
        BufferIterator<AccT> new_centroids_begin, BufferIterator<AccT> new_centroids_end,
        BufferIterator<LabelT> labels_begin, BufferIterator<LabelT> labels_end,
        BufferIterator<MassT> masses_begin, BufferIterator<MassT> masses_end,
        Measurement::DataPoint &datapoint, const WaitList &events)>;

    // One whole single_stage iteration (zero, fused pass, FP64 divide) as one
    // operator: clk_lloyd_step.  sums/masses are outputs; max_shift optional.
    using LloydStepFunction = std::function<Event(
        const Queue &queue, size_t num_features, size_t num_points, size_t num_clusters,
        const PointT *points, const PointT *old_centroids, PointT *new_centroids, AccT *sums,
        LabelT *labels, MassT *masses, double *max_shift, Measurement::DataPoint &datapoint,
        const WaitList &events)>;

    LloydStepFunction create_step(ContextHandle context, FusedConfiguration config,
                                  Measurement::Measurement &measurement)
    {
        (void)create(context, config, measurement);  // validates the strategy, logs parameters
        int id = clk_strategy_id(CLK_STAGE_FUSED, config.strategy.c_str());
        clk_kernel_config k;
        check_status(clk_default_config(context.get(), CLK_STAGE_FUSED, id, &k), "clk_default_config");
        fill_sizes(k, config.global_size, config.local_size);
        k.vector_length = config.vector_length;
        const std::string name = fused_kernel_name(id);
        return [context, k, name](const Queue &q, size_t d, size_t n, size_t kk, const PointT *pts,
                                  const PointT *oc, PointT *nc, AccT *sums, LabelT *labels, MassT *masses,
                                  double *shift, Measurement::DataPoint &dp, const WaitList &events) {
            dp.set_name(name);
            return timed_enqueue(q, dp, events, [&](hipStream_t s) {
                check_status(abi::lloyd_step(context.get(), &k, d, n, kk, pts, oc, nc, sums, labels,
                                            masses, shift, s),
                             "clk_lloyd_step");
            });
        };
    }

    FusedFunction create(ContextHandle context, FusedConfiguration config,
                         Measurement::Measurement &measurement)
    {
        require_supported_types<PointT, LabelT, MassT>();
        measurement.set_parameter("FusedGlobalSize", std::to_string(config.global_size[0]));
        measurement.set_parameter("FusedLocalSize", std::to_string(config.local_size[0]));
        measurement.set_parameter("FusedVectorLength", std::to_string(config.vector_length));
        int id = clk_strategy_id(CLK_STAGE_FUSED, config.strategy.c_str());
        if (id < 0)
            throw std::invalid_argument(config.strategy);
        clk_kernel_config k;
        check_status(clk_default_config(context.get(), CLK_STAGE_FUSED, id, &k), "clk_default_config");
        fill_sizes(k, config.global_size, config.local_size);
        k.vector_length = config.vector_length;
        const std::string name = fused_kernel_name(id);
        return [context, k, name](const Queue &q, size_t d, size_t n, size_t kk,
                                  BufferIterator<PointT> pb, BufferIterator<PointT> pe,
                                  BufferIterator<PointT> ob, BufferIterator<PointT> oe,
                                  BufferIterator<AccT> nb, BufferIterator<AccT> ne,
                                  BufferIterator<LabelT> lb, BufferIterator<LabelT> le,
                                  BufferIterator<MassT> mb, BufferIterator<MassT> me,
                                  Measurement::DataPoint &dp, const WaitList &events) {
            require_len(pb, pe, n * d, "fused points");
            require_len(ob, oe, kk * d, "fused old centroids");
            require_len(nb, ne, kk * d, "fused new centroids");
            require_len(lb, le, n, "fused labels");
            require_len(mb, me, kk, "fused masses");
            dp.set_name(name);
            return timed_enqueue(q, dp, events, [&](hipStream_t s) {
                check_status(abi::fused(context.get(), &k, d, n, kk, pb.get(), ob.get(), nb.get(),
                                       lb.get(), mb.get(), s),
                             "clk_fused");
            });
        };
    }
};

// ------------------------------------------------------------- labeling --

template <typename PointT, typename LabelT, bool ColMajor = true>
class LabelingFactory {
public:
    using LabelingFunction = std::function<Event(
        const Queue &queue, size_t num_features, size_t num_points, size_t num_clusters,
        BufferIterator<PointT> points_begin, BufferIterator<PointT> points_end,
        BufferIterator<PointT> centroids_begin, BufferIterator<PointT> centroids_end,
        BufferIterator<LabelT> labels_begin, BufferIterator<LabelT> labels_end,
        Measurement::DataPoint &datapoint, const WaitList &events)>;

    LabelingFunction create(ContextHandle context, LabelingConfiguration config,
                            Measurement::Measurement &measurement)
    {
        require_supported_types<PointT, LabelT, LabelT>();
        measurement.set_parameter("LabelingGlobalSize", std::to_string(config.global_size[0]));
        measurement.set_parameter("LabelingLocalSize", std::to_string(config.local_size[0]));
        int id = clk_strategy_id(CLK_STAGE_LABELING, config.strategy.c_str());
        if (id < 0)
            throw std::invalid_argument(config.strategy);
        measurement.set_parameter("LabelingVectorLength", std::to_string(config.vector_length));
        measurement.set_parameter("LabelingUnrollClustersLength",
                                  std::to_string(config.unroll_clusters_length));
        measurement.set_parameter("LabelingUnrollFeaturesLength",
                                  std::to_string(config.unroll_features_length));
        clk_kernel_config k;
        check_status(clk_default_config(context.get(), CLK_STAGE_LABELING, id, &k), "clk_default_config");
        fill_sizes(k, config.global_size, config.local_size);
        k.vector_length = config.vector_length;
        return [context, k](const Queue &q, size_t d, size_t n, size_t kk, BufferIterator<PointT> pb,
                            BufferIterator<PointT> pe, BufferIterator<PointT> cb,
                            BufferIterator<PointT> ce, BufferIterator<LabelT> lb,
                            BufferIterator<LabelT> le, Measurement::DataPoint &dp,
                            const WaitList &events) {
            require_len(pb, pe, n * d, "labeling points");
            require_len(cb, ce, kk * d, "labeling centroids");
            require_len(lb, le, n, "labeling labels");
            dp.set_name(k.strategy == CLK_LABELING_UNROLL_VECTOR           ? "LabelingUnrollVector"
                        : k.strategy == CLK_LABELING_UNROLL_VECTOR_BOUNDED ? "LabelingUnrollVectorBounded"
                                                                           : "LabelingMfma");
            return timed_enqueue(q, dp, events, [&](hipStream_t s) {
                check_status(abi::labeling(context.get(), &k, d, n, kk, pb.get(), cb.get(), lb.get(), s),
                             "clk_labeling");
            });
        };
    }
};

// ---------------------------------------------------------- mass update --

template <typename LabelT, typename MassT>
class MassUpdateFactory {
public:
    using MassUpdateFunction = std::function<Event(
        const Queue &queue, size_t num_points, size_t num_clusters,
        BufferIterator<LabelT> labels_begin, BufferIterator<LabelT> labels_end,
        BufferIterator<MassT> masses_begin, BufferIterator<MassT> masses_end,
        Measurement::DataPoint &datapoint, const WaitList &events)>;

    MassUpdateFunction create(ContextHandle context, MassUpdateConfiguration config,
                              Measurement::Measurement &measurement)
    {
        require_supported_types<PointOfLabel<LabelT>, LabelT, MassT>();
        measurement.set_parameter("MassUpdateGlobalSize", std::to_string(config.global_size[0]));
        measurement.set_parameter("MassUpdateLocalSize", std::to_string(config.local_size[0]));
        if (config.strategy == "part_private")
            measurement.set_parameter("MassUpdateVectorLength", std::to_string(config.vector_length));
        int id = clk_strategy_id(CLK_STAGE_MASS_UPDATE, config.strategy.c_str());
        if (id < 0)
            throw std::invalid_argument(config.strategy);
        clk_kernel_config k;
        check_status(clk_default_config(context.get(), CLK_STAGE_MASS_UPDATE, id, &k),
                     "clk_default_config");
        fill_sizes(k, config.global_size, config.local_size);
        k.vector_length = config.vector_length;
        static const char *names[] = {"MassUpdateGlobalAtomic", "MassUpdatePartGlobal",
                                      "MassUpdatePartLocal", "MassUpdatePartPrivate"};
        const std::string name = names[id];
        return [context, k, name](const Queue &q, size_t n, size_t kk, BufferIterator<LabelT> lb,
                                  BufferIterator<LabelT> le, BufferIterator<MassT> mb,
                                  BufferIterator<MassT> me, Measurement::DataPoint &dp,
                                  const WaitList &events) {
            require_len(lb, le, n, "mass_update labels");
            require_len(mb, me, kk, "mass_update masses");
            dp.set_name(name);
            return timed_enqueue(q, dp, events, [&](hipStream_t s) {
                check_status(abi::mass_update(context.get(), &k, n, kk, lb.get(), mb.get(), s),
                             "clk_mass_update");
            });
        };
    }
};

// ------------------------------------------------------ centroid update --

template <typename PointT, typename LabelT, typename MassT, bool ColMajor = true>
class CentroidUpdateFactory {
public:
    using CentroidUpdateFunction = std::function<Event(
        const Queue &queue, size_t num_features, size_t num_points, size_t num_clusters,
        BufferIterator<PointT> points_begin, BufferIterator<PointT> points_end,
        BufferIterator<AccT> centroids_begin, BufferIterator<AccT> centroids_end,
        BufferIterator<LabelT> labels_begin, BufferIterator<LabelT> labels_end,
        BufferIterator<MassT> masses_begin, BufferIterator<MassT> masses_end,
        Measurement::DataPoint &datapoint, const WaitList &events)>;

    CentroidUpdateFunction create(ContextHandle context, CentroidUpdateConfiguration config,
                                  Measurement::Measurement &measurement)
    {
        require_supported_types<PointT, LabelT, MassT>();
        measurement.set_parameter("CentroidUpdateGlobalSize", std::to_string(config.global_size[0]));
        measurement.set_parameter("CentroidUpdateLocalSize", std::to_string(config.local_size[0]));
        if (config.strategy == "cluster_merge" || config.strategy == "feature_sum_pardim")
            measurement.set_parameter("CentroidUpdateVectorLength",
                                      std::to_string(config.vector_length));
        int id = clk_strategy_id(CLK_STAGE_CENTROID_UPDATE, config.strategy.c_str());
        if (id < 0)
            throw std::invalid_argument(config.strategy);
        if (id == CLK_CENTROID_FEATURE_SUM_PARDIM) {
            measurement.set_parameter("CentroidUpdateLocalFeatures", std::to_string(config.local_features));
            measurement.set_parameter("CentroidUpdateThreadFeatures",
                                      std::to_string(config.thread_features));
        }
        clk_kernel_config k;
        check_status(clk_default_config(context.get(), CLK_STAGE_CENTROID_UPDATE, id, &k),
                     "clk_default_config");
        fill_sizes(k, config.global_size, config.local_size);
        k.vector_length = config.vector_length;
        k.local_features = config.local_features;
        k.thread_features = config.thread_features;
        static const char *names[] = {"CentroidUpdateFeatureSum", "CentroidUpdateFeatureSumPardim",
                                      "CentroidUpdateClusterMerge"};
        const std::string name = names[id];
        return [context, k, name](const Queue &q, size_t d, size_t n, size_t kk,
                                  BufferIterator<PointT> pb, BufferIterator<PointT> pe,
                                  BufferIterator<AccT> cb, BufferIterator<AccT> ce,
                                  BufferIterator<LabelT> lb, BufferIterator<LabelT> le,
                                  BufferIterator<MassT> mb, BufferIterator<MassT> me,
                                  Measurement::DataPoint &dp, const WaitList &events) {
            require_len(pb, pe, n * d, "centroid_update points");
            require_len(cb, ce, kk * d, "centroid_update centroids");
            require_len(lb, le, n, "centroid_update labels");
            require_len(mb, me, kk, "centroid_update masses");
            dp.set_name(name);
            return timed_enqueue(q, dp, events, [&](hipStream_t s) {
                check_status(abi::centroid_update(context.get(), &k, d, n, kk, pb.get(), cb.get(),
                                                 lb.get(), mb.get(), s),
                             "clk_centroid_update");
            });
        };
    }
};

// ------------------------------------------------------ MatrixBinaryOp --

// Divide: centroids[f,c] = (PointT)(sums[f,c] / masses[c]) in FP64 (empty
// cluster -> NaN); Add: fst += snd.  The reference divides PointT sums in
// place; here the FP64 sums and the FP32 centroids are separate buffers.
template <typename T1, typename T2>
class MatrixBinaryOp {
public:
    enum BinaryOp { Add, Subtract, Multiply, Divide };

    void prepare(ContextHandle context, BinaryOp op)
    {
        context_ = context;
        op_ = op;
    }

    Event row(const Queue &q, size_t num_cols, size_t num_rows, BufferIterator<AccT> sums_begin,
              BufferIterator<AccT> sums_end, BufferIterator<T2> vector_begin,
              BufferIterator<T2> vector_end, BufferIterator<T1> out_begin,
              BufferIterator<T1> out_end, Measurement::DataPoint &dp, const WaitList &events,
              const T1 *old_centroids = nullptr, double *max_shift = nullptr)
    {
        if (op_ != Divide)
            throw std::invalid_argument("MatrixBinaryOp::row implements Divide");
        require_len(sums_begin, sums_end, num_cols * num_rows, "divide sums");
        require_len(vector_begin, vector_end, num_rows, "divide masses");
        require_len(out_begin, out_end, num_cols * num_rows, "divide centroids");
        dp.set_name("MatrixBinaryOpRow");
        return timed_enqueue(q, dp, events, [&](hipStream_t s) {
            check_status(abi::divide(context_.get(), num_cols, num_rows, sums_begin.get(),
                                               vector_begin.get(), out_begin.get(), old_centroids,
                                               max_shift, s),
                         "clk_matrix_divide_row");
        });
    }

    Event matrix(const Queue &q, size_t num_cols, size_t num_rows, BufferIterator<T1> fst_begin,
                 BufferIterator<T1> fst_end, BufferIterator<T2> snd_begin,
                 BufferIterator<T2> snd_end, Measurement::DataPoint &dp, const WaitList &events)
    {
        static_assert(std::is_same<T1, T2>::value, "matrix() adds equal types");
        if (op_ != Add)
            throw std::invalid_argument("MatrixBinaryOp::matrix implements Add");
        require_len(fst_begin, fst_end, num_cols * num_rows, "matrix fst");
        require_len(snd_begin, snd_end, num_cols * num_rows, "matrix snd");
        dp.set_name("MatrixBinaryOpMatrix");
        return timed_enqueue(q, dp, events, [&](hipStream_t s) {
            check_status(abi::matrix_add(context_.get(), num_cols * num_rows, fst_begin.get(), snd_begin.get(), s),
                         "clk_matrix_add");
        });
    }

private:
    ContextHandle context_;
    BinaryOp op_ = Add;
};

// --------------------------------------------------- ReduceVectorParcol --

template <typename T>
class ReduceVectorParcol {
public:
    void prepare(ContextHandle context) { context_ = context; }

    Event operator()(const Queue &q, size_t num_cols, size_t result_rows, BufferIterator<T> begin,
                     BufferIterator<T> end, Measurement::DataPoint &dp, const WaitList &events)
    {
        if ((size_t)(end - begin) < num_cols * result_rows)
            throw std::length_error("ReduceVectorParcol: buffer smaller than cols * rows");
        dp.set_name("ReduceVectorParcol");
        return timed_enqueue(q, dp, events, [&](hipStream_t s) {
            int rc;
            if constexpr (std::is_same<T, double>::value)
                rc = clk_reduce_vector_parcol_f64(context_.get(), num_cols, result_rows, begin.get(), s);
            else
                rc = clk_reduce_vector_parcol_u32(context_.get(), num_cols, result_rows, begin.get(), s);
            check_status(rc, "clk_reduce_vector_parcol");
        });
    }

private:
    ContextHandle context_;
};

}  // namespace Clustering
