//! `HIP(...)` arguments: the CPU Independent arguments
//! (rustcoalescence/algorithms/independent/src/arguments.rs:79-146, Monolithic
//! parallelism) plus the device form and the launch knobs of nh_job_params.
use necsim_core_bond::{NonNegativeF64, PositiveF64};
use necsim_impls_no_std::parallelisation::independent::{DedupCache, EventSlice, RelativeCapacity};
use necsim_partitioning_core::partition::PartitionSize;
use serde_state::DeserializeState;

use crate::ffi;

/// How the water-level rounds run on the device (nh_job_params.form).
#[derive(Clone, Copy, Debug, PartialEq, Eq, serde::Serialize, serde::Deserialize)]
pub enum HipForm {
    /// One launch per step slice in the reference's queue order
    /// (bit-exact with every DedupCache).
    Plain,
    /// The persistent lane-refilling kernel per round (bit-exact with
    /// DedupCache::None; species and unique events otherwise).
    Persistent,
    /// The CUDA algorithm's task-list launch structure.
    TaskList,
}

#[derive(Clone, Debug, serde::Serialize, serde::Deserialize)]
#[serde(default, deny_unknown_fields)]
pub struct HipArguments {
    pub delta_t: PositiveF64,
    pub step_slice: core::num::NonZeroU64,
    pub dedup_cache: DedupCache,
    pub event_slice: EventSlice,
    pub form: HipForm,
    pub device: i32,
    /// threads per block, 0 = tuned default
    pub block_size: u32,
    /// persistent form: resident blocks per CU, 0 = occupancy limit
    pub blocks_per_cu: u32,
    /// persistent form: capture each round's step / sort / compaction in a hipGraph
    pub use_graph: bool,
    /// persistent form: iterations per lane between refills, 0 = 16
    pub inner_steps: u32,
    /// persistent form: eighths of a wave's active lanes that must hold their
    /// next event before it executes them, 0 = 5
    pub exec_fill: u32,
    /// persistent form: cooperative next-event search, 0 = automatic, 1 = off, 2 = on
    pub coop_search: u32,
    /// plain form: passes per host synchronisation, 0 = 16, 1 = one per pass
    pub plain_batch: u32,
    /// task-list form: lineages per launch, 0 = 4096
    pub task_list_size: u32,
    /// persistent form without dedup: lineages per wave in the pooled
    /// kernel's LDS pool (128 or 192), 0 = one lineage per lane
    pub pool: u32,
    /// persistent form: pick block size, residency, inner steps and fill with
    /// nh_autotune on every 8th lineage over this pause window (0 = off)
    pub autotune_window: f64,
}

impl Default for HipArguments {
    fn default() -> Self {
        Self {
            delta_t: PositiveF64::new(2.0).unwrap(),
            step_slice: core::num::NonZeroU64::new(10).unwrap(),
            dedup_cache: DedupCache::Relative(RelativeCapacity {
                factor: PositiveF64::new(1.0).unwrap(),
            }),
            event_slice: EventSlice::Relative(RelativeCapacity {
                factor: PositiveF64::new(2.0).unwrap(),
            }),
            form: HipForm::Persistent,
            device: 0,
            block_size: 0,
            blocks_per_cu: 0,
            use_graph: false,
            inner_steps: 0,
            exec_fill: 0,
            coop_search: 0,
            plain_batch: 0,
            task_list_size: 0,
            pool: 0,
            autotune_window: 0.0,
        }
    }
}

/// The algorithm enum deserialises its arguments with the partition size as
/// state (rustcoalescence/src/args/config/algorithm.rs:14-34): the device runs
/// the Monolithic parallelism mode, so any other partitioning is refused,
/// like the CUDA algorithm does (cuda/src/arguments.rs:72-107).
impl<'de> DeserializeState<'de, PartitionSize> for HipArguments {
    fn deserialize_state<D>(partition_size: &mut PartitionSize, deserializer: D) -> Result<Self, D::Error>
    where
        D: serde::de::Deserializer<'de>,
    {
        let args = <Self as serde::Deserialize>::deserialize(deserializer)?;
        if !partition_size.is_monolithic() {
            return Err(serde::de::Error::custom(
                "The HIP algorithm runs the monolithic water-level loop on one device and is incompatible with \
                 non-monolithic partitioning.",
            ));
        }
        Ok(args)
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
            form: match self.form {
                HipForm::Plain => ffi::NH_FORM_PLAIN,
                HipForm::Persistent => ffi::NH_FORM_PERSISTENT,
                HipForm::TaskList => ffi::NH_FORM_TASK_LIST,
            },
            block_size: self.block_size,
            blocks_per_cu: self.blocks_per_cu,
            use_graph: u32::from(self.use_graph),
            inner_steps: self.inner_steps,
            exec_fill: self.exec_fill,
            coop_search: self.coop_search,
            plain_batch: self.plain_batch,
            task_list_size: self.task_list_size,
            pool: self.pool,
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
