// pimc_direct.hip -- user-mode AQL dispatch (host code only).
//
// One HSA queue per device, shared by the library's contexts and created on
// first use; each context owns a ring of kernel-argument blocks in device
// memory (a serving slot's ring comes with a queue of its own, so the slot's
// packets run in order without holding up the others).  A dispatch copies the
// arguments through the PCIe BAR, flushes the host data path (HDP) and reads
// the flush register back -- a non-posted read that returns only after the
// writes before it have reached the device -- then writes the packet body, its
// header last, and rings the doorbell.  The packet's agent-scope acquire
// invalidates the scalar cache and the non-coherent cache lines, so a slot
// rewritten through the BAR is read afresh.
#include "pimc_direct.hpp"

#include <dlfcn.h>
#include <hsa/hsa.h>
#include <hsa/hsa_ext_amd.h>
#include <hsa/hsa_ven_amd_loader.h>
#include <immintrin.h>
#include <limits.h>
#include <stdlib.h>

#include <cstring>
#include <mutex>
#include <string>

namespace pimc {
namespace {

constexpr int kMaxDevices = 64;
constexpr uint32_t kQueueSize = 1024;

struct DeviceQueue {
    bool tried = false, ok = false;
    hsa_agent_t gpu{}, cpu{};
    hsa_amd_memory_pool_t vram{};
    hsa_queue_t* q = nullptr;
    volatile uint32_t* hdp = nullptr;
    hsa_signal_t drain{};
    std::mutex m;   // packet slots, doorbell and the drain signal
};

std::mutex g_init;
DeviceQueue g_dev[kMaxDevices];
hsa_ven_amd_loader_1_03_pfn_t g_loader{};
bool g_loader_ok = false;

struct AgentSearch {
    uint32_t bdf = 0, domain = 0;
    hsa_agent_t gpu{}, cpu{};
    bool g = false, c = false;
};

hsa_status_t find_agents(hsa_agent_t a, void* d) {
    AgentSearch* s = static_cast<AgentSearch*>(d);
    hsa_device_type_t t;
    if (hsa_agent_get_info(a, HSA_AGENT_INFO_DEVICE, &t) != HSA_STATUS_SUCCESS) return HSA_STATUS_SUCCESS;
    if (t == HSA_DEVICE_TYPE_CPU && !s->c) {
        s->cpu = a;
        s->c = true;
    }
    if (t == HSA_DEVICE_TYPE_GPU && !s->g) {
        uint32_t bdf = 0, dom = 0;
        if (hsa_agent_get_info(a, (hsa_agent_info_t)HSA_AMD_AGENT_INFO_BDFID, &bdf) == HSA_STATUS_SUCCESS &&
            hsa_agent_get_info(a, (hsa_agent_info_t)HSA_AMD_AGENT_INFO_DOMAIN, &dom) == HSA_STATUS_SUCCESS &&
            (bdf >> 3) == (s->bdf >> 3) && dom == s->domain) {
            s->gpu = a;
            s->g = true;
        }
    }
    return HSA_STATUS_SUCCESS;
}

hsa_status_t find_vram(hsa_amd_memory_pool_t p, void* d) {
    DeviceQueue* D = static_cast<DeviceQueue*>(d);
    hsa_amd_segment_t seg;
    uint32_t flags = 0;
    bool alloc = false;
    if (hsa_amd_memory_pool_get_info(p, HSA_AMD_MEMORY_POOL_INFO_SEGMENT, &seg) != HSA_STATUS_SUCCESS ||
        seg != HSA_AMD_SEGMENT_GLOBAL)
        return HSA_STATUS_SUCCESS;
    if (hsa_amd_memory_pool_get_info(p, HSA_AMD_MEMORY_POOL_INFO_GLOBAL_FLAGS, &flags) != HSA_STATUS_SUCCESS ||
        hsa_amd_memory_pool_get_info(p, HSA_AMD_MEMORY_POOL_INFO_RUNTIME_ALLOC_ALLOWED, &alloc) != HSA_STATUS_SUCCESS)
        return HSA_STATUS_SUCCESS;
    if ((flags & HSA_AMD_MEMORY_POOL_GLOBAL_FLAG_COARSE_GRAINED) && alloc) {
        D->vram = p;
        return HSA_STATUS_INFO_BREAK;
    }
    return HSA_STATUS_SUCCESS;
}

bool setup(int device, DeviceQueue& D) {
    int bus = 0, dev = 0, dom = 0;
    if (hipDeviceGetAttribute(&bus, hipDeviceAttributePciBusId, device) != hipSuccess ||
        hipDeviceGetAttribute(&dev, hipDeviceAttributePciDeviceId, device) != hipSuccess ||
        hipDeviceGetAttribute(&dom, hipDeviceAttributePciDomainId, device) != hipSuccess)
        return false;
    // HIP has initialised the runtime; this takes one more reference on it
    if (hsa_init() != HSA_STATUS_SUCCESS) return false;
    if (!g_loader_ok) {
        if (hsa_system_get_major_extension_table(HSA_EXTENSION_AMD_LOADER, 1, sizeof(g_loader), &g_loader) !=
                HSA_STATUS_SUCCESS ||
            !g_loader.hsa_ven_amd_loader_iterate_executables ||
            !g_loader.hsa_ven_amd_loader_executable_iterate_loaded_code_objects ||
            !g_loader.hsa_ven_amd_loader_loaded_code_object_get_info)
            return false;
        g_loader_ok = true;
    }
    AgentSearch s;
    s.bdf = ((uint32_t)bus << 8) | ((uint32_t)dev << 3);
    s.domain = (uint32_t)dom;
    if (hsa_iterate_agents(find_agents, &s) != HSA_STATUS_SUCCESS || !s.g || !s.c) return false;
    D.gpu = s.gpu;
    D.cpu = s.cpu;
    if (hsa_amd_agent_iterate_memory_pools(D.gpu, find_vram, &D) != HSA_STATUS_INFO_BREAK) return false;
    hsa_amd_hdp_flush_t hdp{};
    if (hsa_agent_get_info(D.gpu, (hsa_agent_info_t)HSA_AMD_AGENT_INFO_HDP_FLUSH, &hdp) != HSA_STATUS_SUCCESS ||
        !hdp.HDP_MEM_FLUSH_CNTL)
        return false;
    D.hdp = hdp.HDP_MEM_FLUSH_CNTL;
    if (hsa_signal_create(0, 0, nullptr, &D.drain) != HSA_STATUS_SUCCESS) return false;
    if (hsa_queue_create(D.gpu, kQueueSize, HSA_QUEUE_TYPE_SINGLE, nullptr, nullptr, UINT32_MAX, UINT32_MAX, &D.q) !=
        HSA_STATUS_SUCCESS) {
        (void)hsa_signal_destroy(D.drain);
        return false;
    }
    return true;
}

DeviceQueue* device_queue(int device) {
    if (device < 0 || device >= kMaxDevices) return nullptr;
    std::lock_guard<std::mutex> lk(g_init);
    DeviceQueue& D = g_dev[device];
    if (!D.tried) {
        D.tried = true;
        D.ok = setup(device, D);
    }
    return D.ok ? &D : nullptr;
}

// Does the loaded code object come from this shared library?  HIP registers
// the code objects embedded in a library either from memory (their bytes
// inside the library's mapping) or from the library file itself.
struct CodeObjectMatch {
    hsa_agent_t agent;
    const void* self_base;
    std::string self_path;
    bool ours = false;
};

std::string uri_decode(const std::string& s) {
    std::string o;
    for (size_t i = 0; i < s.size(); ++i) {
        if (s[i] == '%' && i + 2 < s.size()) {
            o += (char)std::strtol(s.substr(i + 1, 2).c_str(), nullptr, 16);
            i += 2;
        } else {
            o += s[i];
        }
    }
    return o;
}

hsa_status_t check_code_object(hsa_executable_t, hsa_loaded_code_object_t lco, void* d) {
    CodeObjectMatch* m = static_cast<CodeObjectMatch*>(d);
    hsa_agent_t a{};
    if (g_loader.hsa_ven_amd_loader_loaded_code_object_get_info(lco, HSA_VEN_AMD_LOADER_LOADED_CODE_OBJECT_INFO_AGENT,
                                                                &a) != HSA_STATUS_SUCCESS ||
        a.handle != m->agent.handle)
        return HSA_STATUS_SUCCESS;
    uint32_t storage = 0;
    if (g_loader.hsa_ven_amd_loader_loaded_code_object_get_info(
            lco, HSA_VEN_AMD_LOADER_LOADED_CODE_OBJECT_INFO_CODE_OBJECT_STORAGE_TYPE, &storage) == HSA_STATUS_SUCCESS &&
        storage == HSA_VEN_AMD_LOADER_CODE_OBJECT_STORAGE_TYPE_MEMORY) {
        uint64_t base = 0;
        Dl_info info{};
        if (g_loader.hsa_ven_amd_loader_loaded_code_object_get_info(
                lco, HSA_VEN_AMD_LOADER_LOADED_CODE_OBJECT_INFO_CODE_OBJECT_STORAGE_MEMORY_BASE, &base) ==
                HSA_STATUS_SUCCESS &&
            dladdr(reinterpret_cast<const void*>(base), &info) && info.dli_fbase == m->self_base) {
            m->ours = true;
            return HSA_STATUS_INFO_BREAK;
        }
    }
    uint32_t len = 0;
    if (g_loader.hsa_ven_amd_loader_loaded_code_object_get_info(
            lco, HSA_VEN_AMD_LOADER_LOADED_CODE_OBJECT_INFO_URI_LENGTH, &len) != HSA_STATUS_SUCCESS ||
        len == 0 || len > 8192)
        return HSA_STATUS_SUCCESS;
    std::string uri(len + 1, '\0');
    if (g_loader.hsa_ven_amd_loader_loaded_code_object_get_info(lco, HSA_VEN_AMD_LOADER_LOADED_CODE_OBJECT_INFO_URI,
                                                                &uri[0]) != HSA_STATUS_SUCCESS)
        return HSA_STATUS_SUCCESS;
    uri.resize(len);
    const std::string fp = "file://";
    if (uri.compare(0, fp.size(), fp) == 0) {
        std::string path = uri_decode(uri.substr(fp.size(), uri.find_first_of("#?") - fp.size()));
        char rp[PATH_MAX];
        if (realpath(path.c_str(), rp) && m->self_path == rp) {
            m->ours = true;
            return HSA_STATUS_INFO_BREAK;
        }
    } else if (uri.compare(0, 9, "memory://") == 0) {
        const size_t o = uri.find("offset=");
        if (o != std::string::npos) {
            const unsigned long long addr = std::strtoull(uri.c_str() + o + 7, nullptr, 0);
            Dl_info info{};
            if (dladdr(reinterpret_cast<const void*>(addr), &info) && info.dli_fbase == m->self_base) {
                m->ours = true;
                return HSA_STATUS_INFO_BREAK;
            }
        }
    }
    return HSA_STATUS_SUCCESS;
}

struct SymbolSearch {
    hsa_agent_t agent;
    std::string names[2];
    const void* self_base;
    std::string self_path;
    hsa_executable_symbol_t sym{};
    bool found = false;
};

hsa_status_t find_symbol(hsa_executable_t e, void* d) {
    SymbolSearch* s = static_cast<SymbolSearch*>(d);
    for (const std::string& n : s->names) {
        hsa_executable_symbol_t sym;
        if (hsa_executable_get_symbol_by_name(e, n.c_str(), &s->agent, &sym) != HSA_STATUS_SUCCESS) continue;
        hsa_symbol_kind_t kind;
        if (hsa_executable_symbol_get_info(sym, HSA_EXECUTABLE_SYMBOL_INFO_TYPE, &kind) != HSA_STATUS_SUCCESS ||
            kind != HSA_SYMBOL_KIND_KERNEL)
            continue;
        CodeObjectMatch m{s->agent, s->self_base, s->self_path};
        (void)g_loader.hsa_ven_amd_loader_executable_iterate_loaded_code_objects(e, check_code_object, &m);
        if (!m.ours) continue;
        s->sym = sym;
        s->found = true;
        return HSA_STATUS_INFO_BREAK;
    }
    return HSA_STATUS_SUCCESS;
}

}  // namespace

namespace {

// A barrier packet on q completes after every packet before it; waits for it.
void drain(hsa_queue_t* q, hsa_signal_t sig) {
    hsa_signal_store_relaxed(sig, 1);
    const uint64_t idx = hsa_queue_add_write_index_relaxed(q, 1);
    while (idx - hsa_queue_load_read_index_scacquire(q) >= q->size) _mm_pause();
    hsa_barrier_and_packet_t* p = static_cast<hsa_barrier_and_packet_t*>(q->base_address) + (idx & (q->size - 1));
    std::memset(reinterpret_cast<char*>(p) + 4, 0, sizeof(*p) - 4);
    p->completion_signal = sig;
    const uint16_t header = (HSA_PACKET_TYPE_BARRIER_AND << HSA_PACKET_HEADER_TYPE) |
                            (1 << HSA_PACKET_HEADER_BARRIER) |
                            (HSA_FENCE_SCOPE_SYSTEM << HSA_PACKET_HEADER_SCACQUIRE_FENCE_SCOPE) |
                            (HSA_FENCE_SCOPE_SYSTEM << HSA_PACKET_HEADER_SCRELEASE_FENCE_SCOPE);
    __atomic_store_n(reinterpret_cast<uint32_t*>(p), (uint32_t)header, __ATOMIC_RELEASE);
    hsa_signal_store_screlease(q->doorbell_signal, (hsa_signal_value_t)idx);
    (void)hsa_signal_wait_scacquire(sig, HSA_SIGNAL_CONDITION_LT, 1, UINT64_MAX, HSA_WAIT_STATE_BLOCKED);
}

void* alloc_visible(DeviceQueue* D, size_t bytes) {
    void* p = nullptr;
    if (hsa_amd_memory_pool_allocate(D->vram, bytes, 0, &p) != HSA_STATUS_SUCCESS) return nullptr;
    if (hsa_amd_agents_allow_access(1, &D->cpu, nullptr, p) != HSA_STATUS_SUCCESS) {
        (void)hsa_amd_memory_pool_free(p);
        return nullptr;
    }
    return p;
}

}  // namespace

bool direct_ring_create(int device, DirectRing* r, bool own_queue) {
    DeviceQueue* D = device_queue(device);
    if (!D) return false;
    // the host writes the argument blocks through the BAR
    void* p = alloc_visible(D, (size_t)kDirectSlots * kDirectSlotBytes);
    if (!p) return false;
    hsa_queue_t* q = nullptr;
    if (own_queue && hsa_queue_create(D->gpu, kQueueSize, HSA_QUEUE_TYPE_SINGLE, nullptr, nullptr, UINT32_MAX,
                                      UINT32_MAX, &q) != HSA_STATUS_SUCCESS) {
        (void)hsa_amd_memory_pool_free(p);
        return false;
    }
    r->base = static_cast<char*>(p);
    r->device = device;
    r->next = 0;
    r->own_queue = q;
    return true;
}

void direct_ring_release(DirectRing* r) {
    if (!r->base) return;
    DeviceQueue& D = g_dev[r->device];
    if (r->own_queue) {
        hsa_queue_t* q = static_cast<hsa_queue_t*>(r->own_queue);
        hsa_signal_t sig;
        if (hsa_signal_create(0, 0, nullptr, &sig) == HSA_STATUS_SUCCESS) {
            drain(q, sig);
            (void)hsa_signal_destroy(sig);
        }
        (void)hsa_queue_destroy(q);
        r->own_queue = nullptr;
    } else {
        std::lock_guard<std::mutex> lk(D.m);
        drain(D.q, D.drain);
    }
    (void)hsa_amd_memory_pool_free(r->base);
    r->base = nullptr;
}

void* direct_alloc_visible(int device, size_t bytes) {
    DeviceQueue* D = device_queue(device);
    return D ? alloc_visible(D, bytes) : nullptr;
}

void direct_free_visible(void* p) {
    if (p) (void)hsa_amd_memory_pool_free(p);
}

bool direct_kernel(int device, const void* host_stub, hipStream_t stream, uint32_t bytes, DirectKernel* k) {
    DeviceQueue* D = device_queue(device);
    if (!D) return false;
    hipFuncAttributes fa;
    if (hipFuncGetAttributes(&fa, host_stub) != hipSuccess) return false;   // loads the code object
    if (fa.localSizeBytes != 0) return false;                               // needs scratch
    const char* name = hipKernelNameRefByPtr(host_stub, stream);
    if (!name) return false;
    Dl_info self{};
    if (!dladdr(reinterpret_cast<const void*>(&direct_kernel), &self) || !self.dli_fname) return false;
    char rp[PATH_MAX];
    SymbolSearch s;
    s.agent = D->gpu;
    s.names[0] = std::string(name) + ".kd";
    s.names[1] = name;
    s.self_base = self.dli_fbase;
    s.self_path = realpath(self.dli_fname, rp) ? rp : self.dli_fname;
    {
        std::lock_guard<std::mutex> lk(g_init);
        (void)g_loader.hsa_ven_amd_loader_iterate_executables(find_symbol, &s);
    }
    if (!s.found) return false;
    uint64_t obj = 0;
    uint32_t karg = 0, group = 0, priv = 0;
    if (hsa_executable_symbol_get_info(s.sym, HSA_EXECUTABLE_SYMBOL_INFO_KERNEL_OBJECT, &obj) != HSA_STATUS_SUCCESS ||
        hsa_executable_symbol_get_info(s.sym, HSA_EXECUTABLE_SYMBOL_INFO_KERNEL_KERNARG_SEGMENT_SIZE, &karg) !=
            HSA_STATUS_SUCCESS ||
        hsa_executable_symbol_get_info(s.sym, HSA_EXECUTABLE_SYMBOL_INFO_KERNEL_GROUP_SEGMENT_SIZE, &group) !=
            HSA_STATUS_SUCCESS ||
        hsa_executable_symbol_get_info(s.sym, HSA_EXECUTABLE_SYMBOL_INFO_KERNEL_PRIVATE_SEGMENT_SIZE, &priv) !=
            HSA_STATUS_SUCCESS)
        return false;
    // the argument block must be exactly what the caller writes (no hidden
    // arguments the packet would have to supply) and fit one ring slot
    if (obj == 0 || karg != bytes || bytes > kDirectSlotBytes || priv != 0) return false;
    k->object = obj;
    k->kernarg_bytes = karg;
    k->group_bytes = group;
    return true;
}

char* direct_stage(DirectRing* r, const ArgPiece* pieces, int n_pieces) {
    char* slot = r->base + (size_t)(r->next++ % kDirectSlots) * kDirectSlotBytes;
    for (int i = 0; i < n_pieces; ++i) std::memcpy(slot + pieces[i].offset, pieces[i].src, pieces[i].bytes);
    return slot;
}

void direct_flush(int device) {
    DeviceQueue& D = g_dev[device];
    _mm_sfence();
    *D.hdp = 1u;
    (void)*D.hdp;   // non-posted: the BAR writes and the flush have reached the device
}

namespace {
void write_packet(hsa_queue_t* q, const DirectKernel& k, const char* kernargs, uint32_t grid_x, uint32_t block_x) {
    const uint64_t idx = hsa_queue_add_write_index_relaxed(q, 1);
    while (idx - hsa_queue_load_read_index_scacquire(q) >= q->size) _mm_pause();
    hsa_kernel_dispatch_packet_t* p = static_cast<hsa_kernel_dispatch_packet_t*>(q->base_address) +
                                      (idx & (q->size - 1));
    p->workgroup_size_x = (uint16_t)block_x;
    p->workgroup_size_y = 1;
    p->workgroup_size_z = 1;
    p->reserved0 = 0;
    p->grid_size_x = grid_x * block_x;
    p->grid_size_y = 1;
    p->grid_size_z = 1;
    p->private_segment_size = 0;
    p->group_segment_size = k.group_bytes;
    p->kernel_object = k.object;
    p->kernarg_address = const_cast<char*>(kernargs);
    p->reserved2 = 0;
    p->completion_signal.handle = 0;
    // agent-scope fences: the kernels publish their results with system-scope
    // stores themselves, and an agent-scope acquire invalidates the caches that
    // could hold data another kernel changed; the barrier bit keeps a queue's
    // packets in order
    const uint16_t header = (HSA_PACKET_TYPE_KERNEL_DISPATCH << HSA_PACKET_HEADER_TYPE) |
                            (1 << HSA_PACKET_HEADER_BARRIER) |
                            (HSA_FENCE_SCOPE_AGENT << HSA_PACKET_HEADER_SCACQUIRE_FENCE_SCOPE) |
                            (HSA_FENCE_SCOPE_AGENT << HSA_PACKET_HEADER_SCRELEASE_FENCE_SCOPE);
    const uint16_t setup = 1 << HSA_KERNEL_DISPATCH_PACKET_SETUP_DIMENSIONS;
    __atomic_store_n(reinterpret_cast<uint32_t*>(p), (uint32_t)header | ((uint32_t)setup << 16), __ATOMIC_RELEASE);
    hsa_signal_store_screlease(q->doorbell_signal, (hsa_signal_value_t)idx);
}
}  // namespace

void direct_submit(DirectRing* r, const DirectKernel& k, const char* kernargs, uint32_t grid_x, uint32_t block_x) {
    if (r->own_queue) {
        write_packet(static_cast<hsa_queue_t*>(r->own_queue), k, kernargs, grid_x, block_x);
        return;
    }
    DeviceQueue& D = g_dev[r->device];
    std::lock_guard<std::mutex> lk(D.m);
    write_packet(D.q, k, kernargs, grid_x, block_x);
}

void direct_dispatch(DirectRing* r, const DirectKernel& k, const ArgPiece* pieces, int n_pieces, uint32_t grid_x,
                     uint32_t block_x) {
    const char* slot = direct_stage(r, pieces, n_pieces);
    direct_flush(r->device);
    direct_submit(r, k, slot, grid_x, block_x);
}

}  // namespace pimc
