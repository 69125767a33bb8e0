// Plan construction for the runtime (internal header).
#pragma once

#include <hip/hip_runtime.h>

#include <memory>
#include <string>
#include <vector>

#include "op2_hip_rt.h"

namespace op2h {

struct DeviceAllocs {
  std::vector<void *> ptrs;
  template <class T> T *upload(const std::vector<T> &v, hipStream_t s);
  template <class T> T *alloc(size_t n);
  ~DeviceAllocs();
};

struct StagedPlanRec {
  std::string signature;
  op2h_staged_plan plan{};
  std::vector<int> ncolblk_h, colstart_h, max_nthrcol_h;
  DeviceAllocs mem;
};

struct ColorPlanRec {
  std::string signature;
  op2h_color_plan plan{};
  std::vector<int> col_offsets_h;
  DeviceAllocs mem;
};

struct GatherPlanRec {
  op2h_gather_plan plan{};
  DeviceAllocs mem;
};

struct OwnerPlanRec {
  std::string signature;
  op2h_owner_plan plan{};
  int n_owned = 0, n_csr = 0;
  bool key_instances = true;  // instances in the application's element order
  DeviceAllocs mem;
};

// Host-side description of the indirect structure of a loop.
struct LoopShape {
  int set_size = 0;
  int nargs = 0;
  std::vector<int> map_slot;  // per arg: staged map slot or -1
  std::vector<int> loc_slot;  // per arg: loc-map slot or -1
  std::vector<int> dat_slot;  // per arg: conflict (dat) slot or -1
  std::vector<bool> conflict; // per arg: INC/RW through a map
  std::vector<op_map> maps;   // per map slot
  std::vector<std::pair<op_map, int>> locs;  // per loc slot: (map, idx)
  std::vector<int> dat_to_size;              // per dat slot
  std::vector<std::pair<op_map, int>> arg_map;  // per arg (map, idx)
};

LoopShape loop_shape(op_set set, int nargs, const op_arg *args);

// Under transparent renumbering (OP_HIP_RENUMBER): the application's index of
// each element of a renumbered set (old_of_new), else nullptr.  Plans that
// sum increments in element order use it to keep the application's order.
const int *old_index_host(op_set set);

// A map's table in the device numbering, which plans index: under
// transparent renumbering the runtime's renumbered copy (the op_map's own
// table keeps the application's numbering), else the op_map's table.
const int *host_map(op_map m);

// Greedy first-fit colouring in element order with 32 colours per pass
// (op_rt_support.cpp:684-754 for elements, :798-884 for blocks).
// Returns colours per item and the number of colours.
int greedy_block_thread_colours(const LoopShape &sh, int lo, int hi,
                                std::vector<std::vector<unsigned>> &work,
                                int *thrcol);
int greedy_block_colours(const LoopShape &sh, int nblocks, const int *offset,
                         const int *nelems,
                         std::vector<std::vector<unsigned>> &work, int *blkcol);

std::unique_ptr<StagedPlanRec> build_staged_plan(op_set set, int part_size,
                                                 int nargs, const op_arg *args,
                                                 int on_device,
                                                 hipStream_t stream);
std::unique_ptr<ColorPlanRec> build_color_plan(op_set set, int nargs,
                                               const op_arg *args,
                                               int on_device,
                                               hipStream_t stream);
std::unique_ptr<GatherPlanRec> build_gather_plan(op_map map, hipStream_t stream);
// Owner-compute partition plan (op2h_owner_plan); nullptr and `err` set when
// the loop does not have the required shape.
// The owner partition's lists for targets [0,nto) of a map (nel x dim,
// row-major) incremented through the map columns inc_idx[0..ninc): blocks of
// block_cells targets (the last may be smaller), per block its owned targets
// (ascending) and the elements touching them (ascending), per owned target
// its increments as 16-bit (local element * ninc + k) entries in element
// order.  False, with err set, when a block's local ids overflow 16 bits.
struct OwnerLists {
  int nblocks = 0, max_elems = 0, max_owned = 0, max_incs = 0;
  std::vector<int> owned_offs, owned, elem_offs, elems, csr_offs;
  std::vector<unsigned short> csr;
};
// elem_key (may be nullptr): the order in which one target's increments are
// listed is ascending (elem_key[element], k) instead of (element, k); with
// key_instances a block's instances follow that order too (else ascending
// element, which keeps neighbouring elements in neighbouring lanes).
bool owner_lists(int nto, int nel, int dim, const int *map, int ninc,
                 const int *inc_idx, int block_cells, OwnerLists &L,
                 std::string &err, const int *elem_key = nullptr,
                 bool key_instances = true);

std::unique_ptr<OwnerPlanRec> build_owner_plan(op_set set, int block_cells,
                                               int nargs, const op_arg *args,
                                               hipStream_t stream,
                                               std::string &err,
                                               bool key_instances = true);
int download_owner(const OwnerPlanRec &r, int *owned_offs, int *owned,
                   int *elem_offs, int *elems, int *csr_offs,
                   unsigned short *csr);

// GPU colouring (op2h_plan_gpu.hip).  gpu_colour_elements colours elements
// [0,n) so that no two elements sharing an increment target get the same
// colour; gpu_colour_blocks_and_threads colours the threads of each plan block
// and the blocks, from the plan's uploaded staging lists.  Both return the
// number of colours, or -1 when the device path does not apply (the caller
// then colours on the host).
int gpu_colour_elements(const LoopShape &sh, int n, int *col_d,
                        hipStream_t stream);
int gpu_colour_blocks_and_threads(const LoopShape &sh, int nblocks,
                                  int part_size, const op2h_staged_plan &P,
                                  int *blkcol_d, unsigned char *thrcol_d,
                                  int *nthrcol_d, hipStream_t stream);

// Plan checks (op2h_plan_check.cpp) of a freshly built plan: the plan is read
// back and checked on the host; on a failure the messages are printed and the
// program exits with -1.
void check_staged_or_exit(const StagedPlanRec &rec, const char *name, op_set set,
                          int nargs, const op_arg *args);
void check_color_or_exit(const ColorPlanRec &rec, const char *name, op_set set,
                         int nargs, const op_arg *args);
void check_owner_or_exit(const OwnerPlanRec &rec, const char *name, int nargs,
                         const op_arg *args);
void check_gather_or_exit(const GatherPlanRec &rec, op_map map);

int download_staged(const op2h_staged_plan *p, int *thrcol, int *blkcol,
                    int *offset, int *nelems);
int download_color(const op2h_color_plan *p, int *color_of_elem);

}  // namespace op2h
