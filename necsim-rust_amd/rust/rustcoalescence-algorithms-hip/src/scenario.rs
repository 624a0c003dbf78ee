//! Scenario cogs -> `nh_scenario` for the three scenario families the device
//! supports (non-spatial, almost-infinite normal, spatially explicit with a
//! packed (separable) alias dispersal sampler).
//!
//! The in-memory tables are read through accessors that a maintainer adds to
//! the reference types, mirroring the `LendRustToCuda` borrow the CUDA
//! algorithm takes of the same fields
//! (.../dispersal_sampler/in_memory/packed_separable_alias/mod.rs:47-58):
//!   InMemoryHabitat::{habitat(), extent()}
//!   InMemoryPackedSeparableAliasDispersalSampler::{alias_dispersal_ranges(),
//!       self_dispersal(), alias_dispersal_buffer()}
use necsim_core::{cogs::MathsCore, lineage::Lineage};
use necsim_impls_no_std::cogs::{origin_sampler::pre_sampler::OriginPreSampler, rng::wyhash::WyHash};
use rustcoalescence_scenarios::{Scenario, ScenarioCogs};

use crate::ffi;

pub trait HipScenario<M: MathsCore>: Scenario<M, WyHash<M>> + Sized {
    /// The device description of the cogs and whatever must outlive it.
    fn hip_scenario(cogs: &ScenarioCogs<M, WyHash<M>, Self>) -> (ffi::nh_scenario, Box<dyn core::any::Any>);
    fn speciation_probability(cogs: &ScenarioCogs<M, WyHash<M>, Self>) -> f64;
    fn turnover_rate(cogs: &ScenarioCogs<M, WyHash<M>, Self>) -> f64;
    /// `O::sample_habitat(&habitat, pre_sampler, aux)` collected in order
    /// (independent/src/launch.rs:78-86, active_lineage_sampler/independent/mod.rs:98-117)
    fn sample_lineages<I: Iterator<Item = u64>>(
        cogs: &ScenarioCogs<M, WyHash<M>, Self>,
        pre_sampler: OriginPreSampler<M, I>,
    ) -> Vec<Lineage>;
}
