// serve_policy.h -- admission and sample allocation of the render server.
//
// Pure host code with no device dependency, so the policy is testable on a
// machine without a GPU (fh_policy_* entry points).
//
// Admission is first-come first-served by HBM footprint: the head of the
// queue is admitted while the footprints of the active requests plus its own
// fit the budget (no overtaking, so no request starves).  A request larger
// than the whole budget is rejected.
//
// Each iteration advances every active request by k_i camera samples, sized so
// the iteration traces about `target_paths` light-sample paths: requests share
// the target equally, each takes at least one sample and at most what it has
// left, and what the small ones leave unused goes to the others in admission
// order.  Small requests therefore finish in few iterations while large ones
// keep the device full.
#pragma once

#include <stdint.h>

#include <algorithm>
#include <deque>
#include <vector>

namespace fh
{
struct PolicyItem
    {
    uint64_t id;
    uint64_t footprint;        // bytes resident while active
    uint64_t paths_per_sample; // W * H * light_samples
    uint32_t remaining;        // camera samples still to render
    };

// Bytes a request keeps resident: scene (bounds, sort keys, nodes, parents,
// primitive data) plus outputs plus the queues and records of one sample.
inline uint64_t request_footprint(uint64_t n_prims, uint64_t prim_bytes, uint64_t W, uint64_t H, uint64_t L)
    {
    // bounds, sort keys, binary node, 4-wide node and its index, parent, flags
    const uint64_t scene = n_prims * (32 + 16 + 64 + 128 + 4 + 8 + 4 + prim_bytes);
    const uint64_t outputs = W * H * (16 + 4 + 16 + 4);
    const uint64_t per_sample = W * H * L * (2 * 64 + 16 + 16) + W * H * 64;
    return scene + outputs + per_sample;
    }

// Bytes of one render scratch for batches of `paths` light-sample paths: two
// ray queues (64 B per path each), closest hits and termination records
// (16 B each), plus `fixed` (overflow stacks, counters) that does not depend
// on the batch size.  The server keeps two such scratches beside the
// requests' footprints, so admission uses the budget minus them.
inline uint64_t scratch_bytes(uint64_t paths, uint64_t fixed)
    {
    return paths * (2 * 64 + 16 + 16) + fixed;
    }

// Admit from the queue front while the budget allows.  Returns the admitted
// items (removed from `queue`) and the rejected ones (larger than the budget).
inline void policy_admit(std::deque<PolicyItem>& queue, uint64_t budget, uint64_t& in_use,
                         std::vector<PolicyItem>& admitted, std::vector<PolicyItem>& rejected)
    {
    while (!queue.empty())
        {
        const PolicyItem& it = queue.front();
        if (it.footprint > budget)
            {
            rejected.push_back(it);
            queue.pop_front();
            continue;
            }
        if (in_use + it.footprint > budget)
            break;
        in_use += it.footprint;
        admitted.push_back(it);
        queue.pop_front();
        }
    }

// Samples per active request for one iteration.
inline std::vector<uint32_t> policy_allocate(const std::vector<PolicyItem>& active, uint64_t target_paths)
    {
    std::vector<uint32_t> k(active.size(), 0);
    if (active.empty())
        return k;
    const uint64_t share = std::max<uint64_t>(1, target_paths / active.size());
    uint64_t used = 0;
    for (size_t i = 0; i < active.size(); i++)
        {
        const uint64_t pps = std::max<uint64_t>(1, active[i].paths_per_sample);
        uint64_t want = std::max<uint64_t>(1, share / pps);
        want = std::min<uint64_t>(want, active[i].remaining);
        k[i] = (uint32_t)want;
        used += want * pps;
        }
    // hand the unused part of the target out in admission order
    for (size_t i = 0; i < active.size() && used < target_paths; i++)
        {
        const uint64_t pps = std::max<uint64_t>(1, active[i].paths_per_sample);
        const uint64_t room = active[i].remaining - k[i];
        const uint64_t more = std::min<uint64_t>(room, (target_paths - used) / pps);
        k[i] += (uint32_t)more;
        used += more * pps;
        }
    return k;
    }
    } // namespace fh
