// topology.hpp -- math-model topology, bus ordering and the device factorisation plan.
//
// Bus numbering must equal the reference's bit for bit (it fixes the fill-in pattern and the FP
// operation order): topology.hpp:196-368 (CSR digraph, depth-first visit per connected source,
// reversed discovery order for trees, minimum degree on the cyclic part) and sparse_ordering.hpp:20-219.
// Y-bus LU pattern: y_bus.hpp:122-293 (entries of all branches, shunts and fill-ins, CSR sorted by column).
#pragma once

#include "model.hpp"

#include <array>
#include <utility>
#include <vector>

namespace pgm_amd {

struct MathModel {
    Idx n_bus = 0;
    Idx slack = 0;
    bool radial = true;                            // no cycle met by the depth-first visit
    std::vector<Idx> bus_node;                     // topological node of each bus
    std::vector<Idx> branch;                       // component branch index of each math branch
    std::vector<std::array<Idx, 2>> branch_bus;    // -1 for an open side
    std::vector<std::array<Idx, 2>> fill_in;       // bus pairs
    std::vector<Idx> load, source, shunt;          // component indices, stable by bus
};

struct Topology {
    std::vector<Idx> node_group, node_pos;
    std::vector<double> node_phase;  // phase shift against the model's source (transformer clocks)
    std::vector<MathModel> models;
    std::vector<Idx> branch_group, branch_pos;
    std::vector<Idx> lg_group, lg_pos, src_group, src_pos, sh_group, sh_pos;
};

// couple_open (per component branch, optional): branches open on both sides in st that still become
// math branches (both sides -1) of the model holding both their nodes; union groups close them per lane
Topology build_topology(Model const& m, ComponentState const& st, std::vector<char> const* couple_open = nullptr);

// minimum degree ordering on an adjacency map (keys sorted) -- returns (order, fills)
std::pair<std::vector<Idx>, std::vector<std::pair<Idx, Idx>>>
minimum_degree(std::vector<Idx> const& keys, std::vector<std::vector<Idx>> const& adjacency);

// Device plan of one group of math models solved together (block diagonal, concatenated buses).
struct Plan {
    int P = 1;          // phases (1 symmetric, 3 asymmetric)
    Idx n = 0;          // buses
    Idx E = 0;          // LU entries
    Idx n_mm = 0;
    std::vector<int> mm_bus_begin;  // n_mm + 1
    // row p's entries: row_ptr[p], row_ptr[p] + stride, ... < row_end[p], stride = row_stride[p] (empty:
    // 1, plain CSR)
    std::vector<int> row_ptr, row_end, row_stride, col, diag, transpose;
    std::vector<int8_t> fill;  // entry is a fill-in
    std::vector<int> bus_mm;    // math model of every bus, -1 for an unused (padding) bus label
    std::vector<double> phase;  // phase shift of every bus
    std::vector<int> mm_src_ptr, mm_src;  // sources of every math model
    // label of every bus in math-model order (mm_bus_begin[k] + position); identity unless interleaved
    std::vector<int> bus_label;
    // Y values per LU entry (P*P complex, row-major), zero for fill-ins
    std::vector<cd> ydata;
    // Schur updates in pull form, per pivot: (target, l, u) entry triples, k ascending per target
    std::vector<int> schur_ptr;
    std::vector<std::array<int, 3>> schur;
    // per pivot: the row has exactly one U entry and every Schur update it owns targets its diagonal
    std::vector<int8_t> keep;
    // tasks: disjoint sets of pivots; a task only depends on tasks of lower level
    std::vector<int> task_ptr, task_pivot, task_mm;
    std::vector<int> level_ptr, level_task;  // tasks per level
    // appliances, concatenated over the math models in math order
    std::vector<int> load_ptr, src_ptr, shunt_ptr;  // per bus
    std::vector<int8_t> load_type;
    std::vector<Idx> load_comp, src_comp, shunt_comp;
    std::vector<cd> shunt_y;  // P*P per shunt
    // branches
    std::vector<int> branch_from, branch_to;  // bus or -1
    std::vector<Idx> branch_comp;
    std::vector<cd> branch_y;  // 4 * P*P per branch
};

// Schur update lists and the task / level schedule of a plan whose pattern (n, E, row_ptr, col, diag,
// transpose, bus_mm) is set
void build_schedule(Plan& p, int target_parallel_tasks);

Plan build_plan(Model const& m, ComponentState const& st, Topology const& topo, std::vector<Idx> const& models,
                bool symmetric, int target_parallel_tasks);

// Relabel buses and LU entries of a plan for waves that run tasks_per_wave tasks of one level side by
// side (packed groups): in a wave chunk of nt tasks the k-th pivot of task t becomes bus B0 + k nt + t
// and its row entries E0 + (k R + j) nt + t, so lanes at the same step touch consecutive addresses.
void interleave_plan(Plan& p, int tasks_per_wave);

// Structural consistency of a plan (rows sorted, one diagonal, transposes paired, Schur triples
// inside the pivot's row / column, task pivots cover the buses); throws std::logic_error
void check_plan(Plan const& p);

}  // namespace pgm_amd
