// launch_sink.h -- kernel launches that are either issued on a stream at once
// or recorded as hipGraph kernel nodes (function, shape and a copy of the
// arguments).  Host C++ only: it needs the HIP runtime API, no device code.
//
// kemit takes the kernel as a typed function pointer and converts every
// argument to the kernel's own parameter type before it is launched or
// copied into a node.  HIP reads exactly sizeof(parameter) bytes through each
// kernelParams pointer, so an argument of another width (an int where the
// kernel takes a long) would otherwise be read past its slot.
#pragma once
#include <hip/hip_runtime_api.h>

#include <cstddef>
#include <cstring>
#include <tuple>
#include <type_traits>
#include <utility>
#include <vector>

namespace txgpu {

struct KNode {
    const void *func = nullptr;
    dim3 grid, block;
    unsigned shm = 0;
    std::vector<unsigned char> blob; // arguments, each at offs[i]
    std::vector<size_t> offs;
    std::vector<size_t> sizes;       // sizeof each argument
    std::vector<void *> params()     // pointers into blob, valid while blob is
    {
        std::vector<void *> v;
        for (size_t o : offs) v.push_back(blob.data() + o);
        return v;
    }
};

// Where the launch helpers put their launches: nullptr = launch on the
// stream now; otherwise appended to nodes (a graph under construction).
struct LaunchSink {
    std::vector<KNode> nodes;
};

template <typename T>
inline void knode_put(KNode &n, const T &v)
{
    static_assert(std::is_trivially_copyable<T>::value, "kernel arguments are copied bytewise");
    const size_t o = (n.blob.size() + alignof(T) - 1) / alignof(T) * alignof(T);
    n.blob.resize(o + sizeof(T));
    std::memcpy(n.blob.data() + o, &v, sizeof(T));
    n.offs.push_back(o);
    n.sizes.push_back(sizeof(T));
}

namespace detail {
template <typename Tuple, size_t... I>
inline hipError_t launch_tuple(const void *f, dim3 g, dim3 b, size_t shm, hipStream_t st, Tuple &t,
                               std::index_sequence<I...>)
{
    void *p[] = {(void *)&std::get<I>(t)..., nullptr};
    return hipLaunchKernel(f, g, b, p, shm, st);
}
} // namespace detail

// Launches f(args...) on st, or records it in sink.  Each argument is
// converted to the matching parameter type of f.
template <typename... P, typename... A>
inline hipError_t kemit(LaunchSink *sink, hipStream_t st, void (*f)(P...), dim3 g, dim3 b, size_t shm,
                        A... args)
{
    static_assert(sizeof...(P) == sizeof...(A), "argument count differs from the kernel's parameters");
    std::tuple<typename std::decay<P>::type...> t(static_cast<typename std::decay<P>::type>(args)...);
    if (!sink)
        return detail::launch_tuple((const void *)f, g, b, shm, st, t, std::index_sequence_for<P...>{});
    KNode n;
    n.func = (const void *)f;
    n.grid = g;
    n.block = b;
    n.shm = (unsigned)shm;
    std::apply([&n](const auto &...v) { (knode_put(n, v), ...); }, t);
    sink->nodes.push_back(std::move(n));
    return hipSuccess;
}

} // namespace txgpu
