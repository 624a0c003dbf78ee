//! `HIP(...)` arguments: the CPU Independent arguments
//! (rustcoalescence/algorithms/independent/src/arguments.rs:79-146) plus the
//! device form and launch knobs.
use necsim_core_bond::{NonNegativeF64, PositiveF64};
use necsim_impls_no_std::parallelisation::independent::{DedupCache, EventSlice};

use crate::ffi;

#[derive(Clone, Debug, serde::Serialize, serde::Deserialize)]
#[serde(default, deny_unknown_fields)]
pub struct HipArguments {
    pub delta_t: PositiveF64,
    pub step_slice: core::num::NonZeroU64,
    pub dedup_cache: DedupCache,
    pub event_slice: EventSlice,
    /// "Plain" (one launch per step slice, reference queue order) or "Persistent"
    pub persistent: bool,
    pub device: i32,
    pub block_size: u32,
    pub blocks_per_cu: u32,
}

impl Default for HipArguments {
    fn default() -> Self {
        Self {
            delta_t: PositiveF64::new(2.0).unwrap(),
            step_slice: core::num::NonZeroU64::new(10).unwrap(),
            dedup_cache: DedupCache::Relative(necsim_impls_no_std::parallelisation::independent::RelativeCapacity {
                factor: PositiveF64::new(1.0).unwrap(),
            }),
            event_slice: EventSlice::Relative(necsim_impls_no_std::parallelisation::independent::RelativeCapacity {
                factor: PositiveF64::new(2.0).unwrap(),
            }),
            persistent: true,
            device: 0,
            block_size: 0,
            blocks_per_cu: 0,
        }
    }
}

impl HipArguments {
    pub(crate) fn job_params(
        &self,
        rng_seed: u64,
        speciation_probability: f64,
        turnover_rate: f64,
        resume_after: Option<NonNegativeF64>,
        pause_before: Option<NonNegativeF64>,
    ) -> ffi::nh_job_params {
        let mut p = ffi::nh_job_params {
            rng_seed,
            speciation_probability,
            turnover_rate,
            delta_t: self.delta_t.get(),
            step_slice: self.step_slice.get(),
            resume_after: resume_after.map_or(0.0, NonNegativeF64::get),
            form: if self.persistent { ffi::NH_FORM_PERSISTENT } else { ffi::NH_FORM_PLAIN },
            block_size: self.block_size,
            blocks_per_cu: self.blocks_per_cu,
            ..Default::default()
        };
        match self.dedup_cache {
            DedupCache::None => p.dedup_mode = ffi::NH_DEDUP_NONE,
            DedupCache::Relative(r) => {
                p.dedup_mode = ffi::NH_DEDUP_RELATIVE;
                p.dedup_factor = r.factor.get();
            },
            DedupCache::Absolute(a) => {
                p.dedup_mode = ffi::NH_DEDUP_ABSOLUTE;
                p.dedup_capacity = a.capacity.get() as u64;
            },
        }
        match self.event_slice {
            EventSlice::Relative(r) => {
                p.event_slice_mode = ffi::NH_SLICE_RELATIVE;
                p.event_slice_factor = r.factor.get();
            },
            EventSlice::Absolute(a) => {
                p.event_slice_mode = ffi::NH_SLICE_ABSOLUTE;
                p.event_slice_capacity = a.capacity.get() as u64;
            },
        }
        if let Some(pause) = pause_before {
            p.has_pause_before = 1;
            p.pause_before = pause.get();
        }
        p
    }
}
