// k-molecule merge (placeholder until the device merge lands).
#include "swap_host.h"

namespace swapk {

int host_merge(Pipeline &p, const swap_cfg &cfg, std::vector<uint32_t> &visit, HostCompacted &out) {
    (void)p; (void)cfg; (void)visit; (void)out;
    return SWAP_E_STATE;
}

}  // namespace swapk
