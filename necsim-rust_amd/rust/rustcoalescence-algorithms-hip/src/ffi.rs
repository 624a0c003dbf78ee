//! Raw bindings of include/necsim_hip.h (hand-written, no bindgen).
#![allow(non_camel_case_types)]

use core::ffi::{c_char, c_int};

pub const NH_OK: c_int = 0;
pub const NH_E_INVALID: c_int = -1;
pub const NH_E_STATE: c_int = -4;

#[repr(C)]
pub struct nh_ctx {
    _private: [u8; 0],
}
#[repr(C)]
pub struct nh_tables {
    _private: [u8; 0],
}
#[repr(C)]
pub struct nh_job {
    _private: [u8; 0],
}

/// Same layout as necsim_core::lineage::Lineage (#[repr(C)], 32 B).
#[repr(C)]
#[derive(Clone, Copy, Default)]
pub struct nh_lineage {
    pub global_reference: u64,
    pub last_event_time: f64,
    pub x: u32,
    pub y: u32,
    pub index: u32,
    pub _pad: u32,
}

/// Same layout as necsim_core::event::SpeciationEvent (#[repr(C)], 40 B).
#[repr(C)]
#[derive(Clone, Copy, Default)]
pub struct nh_speciation_event {
    pub global_reference: u64,
    pub prior_time: f64,
    pub event_time: f64,
    pub x: u32,
    pub y: u32,
    pub index: u32,
    pub _pad: u32,
}

/// Same layout as necsim_core::event::PackedEvent (#[repr(C)], 56 B).
#[repr(C)]
#[derive(Clone, Copy, Default)]
pub struct nh_packed_event {
    pub global_reference: u64,
    pub prior_time: f64,
    pub event_time: f64,
    pub origin_x: u32,
    pub origin_y: u32,
    pub origin_index: u32,
    pub target_x: u32,
    pub target_y: u32,
    pub target_index: u32,
    pub coalescence: u64,
}

/// Same layout as necsim_impls_no_std::alias::packed::AliasMethodSamplerAtom<usize>.
#[repr(C)]
#[derive(Clone, Copy, Default)]
pub struct nh_alias_atom {
    pub u: f64,
    pub e: u64,
    pub k: u64,
}
/// Same layout as .../in_memory/packed_alias::AliasSamplerRange.
#[repr(C)]
#[derive(Clone, Copy, Default)]
pub struct nh_alias_range {
    pub start: u64,
    pub end: u64,
}
/// Same layout as .../in_memory/packed_separable_alias::SeparableAliasSelfDispersal.
#[repr(C)]
#[derive(Clone, Copy, Default)]
pub struct nh_self_dispersal {
    pub self_dispersal: f64,
    pub non_self_dispersal_event: u64,
}

pub const NH_SCENARIO_NON_SPATIAL: i32 = 0;
pub const NH_SCENARIO_ALMOST_INFINITE_NORMAL: i32 = 1;
pub const NH_SCENARIO_SPATIALLY_EXPLICIT: i32 = 2;
pub const NH_SCENARIO_SPATIALLY_EXPLICIT_PACKED: i32 = 3;

#[repr(C)]
#[derive(Clone, Copy)]
pub struct nh_scenario {
    pub kind: i32,
    pub width: u32,
    pub height: u32,
    pub origin_x: u32,
    pub origin_y: u32,
    pub deme: u32,
    pub sigma: f64,
    pub habitat: *const u32,
    pub alias_ranges: *const nh_alias_range,
    pub self_dispersal: *const nh_self_dispersal,
    pub atoms: *const nh_alias_atom,
    pub n_atoms: usize,
}

impl Default for nh_scenario {
    fn default() -> Self {
        Self {
            kind: NH_SCENARIO_NON_SPATIAL,
            width: 0,
            height: 0,
            origin_x: 0,
            origin_y: 0,
            deme: 0,
            sigma: 0.0,
            habitat: core::ptr::null(),
            alias_ranges: core::ptr::null(),
            self_dispersal: core::ptr::null(),
            atoms: core::ptr::null(),
            n_atoms: 0,
        }
    }
}

pub const NH_DEDUP_NONE: i32 = 0;
pub const NH_DEDUP_RELATIVE: i32 = 1;
pub const NH_DEDUP_ABSOLUTE: i32 = 2;
pub const NH_SLICE_RELATIVE: i32 = 1;
pub const NH_SLICE_ABSOLUTE: i32 = 2;
pub const NH_FORM_PLAIN: i32 = 0;
pub const NH_FORM_PERSISTENT: i32 = 1;
pub const NH_FORM_TASK_LIST: i32 = 2;

pub const NH_LINEAGE_OK: i32 = 0;
pub const NH_LINEAGE_OUT_OF_DEME: i32 = 1;
pub const NH_LINEAGE_OUT_OF_HABITAT: i32 = 2;

#[repr(C)]
#[derive(Clone, Copy, Default)]
pub struct nh_job_params {
    pub rng_seed: u64,
    pub speciation_probability: f64,
    pub turnover_rate: f64,
    pub delta_t: f64,
    pub step_slice: u64,
    pub dedup_mode: i32,
    pub dedup_factor: f64,
    pub dedup_capacity: u64,
    pub event_slice_mode: i32,
    pub event_slice_factor: f64,
    pub event_slice_capacity: u64,
    pub has_pause_before: i32,
    pub pause_before: f64,
    pub resume_after: f64,
    pub form: i32,
    pub block_size: u32,
    pub blocks_per_cu: u32,
    pub use_graph: u32,
    pub inner_steps: u32,
    pub exec_fill: u32,
    pub report_dispersal: i32,
    pub event_buffer: u64,
    pub device_event_counter: i32,
    pub plain_batch: u32,
    pub coop_search: u32,
    pub task_list_size: u32,
    pub pool: u32,
}

#[repr(C)]
#[derive(Clone, Copy, Default)]
pub struct nh_outcome {
    pub paused: i32,
    pub time: f64,
    pub steps: u64,
    pub n_events: u64,
    pub n_remaining: u64,
    pub rounds: u64,
    pub launches: u64,
}

extern "C" {
    pub fn nh_last_error(buf: *mut c_char, cap: usize) -> c_int;
    pub fn nh_init(device: c_int, out: *mut *mut nh_ctx) -> c_int;
    pub fn nh_ctx_destroy(ctx: *mut nh_ctx) -> c_int;
    pub fn nh_tables_create(ctx: *mut nh_ctx, s: *const nh_scenario, out: *mut *mut nh_tables) -> c_int;
    pub fn nh_tables_destroy(t: *mut nh_tables) -> c_int;
    pub fn nh_classify_lineages(
        t: *const nh_tables,
        l: *const nh_lineage,
        n: usize,
        kinds: *mut i32,
        n_exceptional: *mut usize,
    ) -> c_int;
    pub fn nh_job_create(ctx: *mut nh_ctx, t: *const nh_tables, p: *const nh_job_params, out: *mut *mut nh_job) -> c_int;
    pub fn nh_job_add_lineages(job: *mut nh_job, l: *const nh_lineage, n: usize) -> c_int;
    pub fn nh_job_add_origins(
        job: *mut nh_job,
        percentage: f64,
        rank: u32,
        size: u32,
        circle_centre_x: u32,
        circle_centre_y: u32,
        circle_radius: u32,
    ) -> c_int;
    pub fn nh_job_start(job: *mut nh_job) -> c_int;
    pub fn nh_job_step_round(job: *mut nh_job, finished: *mut i32) -> c_int;
    pub fn nh_job_drain_events(job: *mut nh_job, buf: *mut nh_speciation_event, cap: usize, n_out: *mut usize) -> c_int;
    pub fn nh_job_drain_packed_events(job: *mut nh_job, buf: *mut nh_packed_event, cap: usize, n_out: *mut usize) -> c_int;
    pub fn nh_job_outcome(job: *const nh_job, out: *mut nh_outcome) -> c_int;
    pub fn nh_job_take_lineages(job: *mut nh_job, buf: *mut nh_lineage, cap: usize, n_out: *mut usize) -> c_int;
    pub fn nh_job_destroy(job: *mut nh_job) -> c_int;
    pub fn nh_autotune(
        ctx: *mut nh_ctx,
        t: *const nh_tables,
        base: *const nh_job_params,
        sample: *const nh_lineage,
        n: usize,
        window: f64,
        tuned: *mut nh_job_params,
        best_ms: *mut f64,
    ) -> c_int;
}
