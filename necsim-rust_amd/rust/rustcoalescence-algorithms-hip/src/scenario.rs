//! Scenario cogs -> `nh_scenario`, the description nh_tables_create uploads.
//!
//! The CUDA algorithm reaches its cogs through `RustToCuda` bounds on each
//! cog type (rustcoalescence/algorithms/cuda/src/lib.rs:79-89).  The device
//! library here implements a fixed set of cogs, so the glue describes them per
//! scenario: every scenario the CLI can be built with implements
//! `HipScenario` (the dispatch in rustcoalescence/src/cli/simulate/dispatch/
//! valid/algorithm_scenario.rs instantiates every algorithm with every
//! scenario), and those whose cogs the device does not implement refuse at
//! run time.
//!
//! The habitat, turnover rate and speciation probability are read through
//! their traits (Habitat::get_extent / get_habitat_at_location,
//! TurnoverRate::get_turnover_rate_at_location,
//! SpeciationProbability::get_speciation_probability_at_location); the
//! dispersal samplers' parameters and tables through the accessors that
//! reference-registration.patch adds, the fields that the CUDA algorithm's
//! `LendRustToCuda` derive borrows
//! (.../dispersal_sampler/in_memory/packed_separable_alias/mod.rs:47-58).
use necsim_core::{
    cogs::{Habitat, MathsCore, SpeciationProbability, TurnoverRate},
    landscape::Location,
};
use necsim_impls_no_std::cogs::rng::wyhash::WyHash;
use rustcoalescence_scenarios::{Scenario, ScenarioCogs};

use crate::{ffi, HipError};

/// The device description of a scenario's cogs together with the host
/// buffers its pointers refer to (they must outlive nh_tables_create).
pub struct HipCogs {
    pub desc: ffi::nh_scenario,
    pub speciation_probability: f64,
    pub turnover_rate: f64,
    habitat: Vec<u32>,
    ranges: Vec<ffi::nh_alias_range>,
    self_dispersal: Vec<ffi::nh_self_dispersal>,
    atoms: Vec<ffi::nh_alias_atom>,
}

impl HipCogs {
    fn new(desc: ffi::nh_scenario, speciation_probability: f64, turnover_rate: f64) -> Self {
        Self {
            desc,
            speciation_probability,
            turnover_rate,
            habitat: Vec::new(),
            ranges: Vec::new(),
            self_dispersal: Vec::new(),
            atoms: Vec::new(),
        }
    }

    /// the description with its pointers set to the owned buffers
    pub fn scenario(&self) -> ffi::nh_scenario {
        ffi::nh_scenario {
            habitat: if self.habitat.is_empty() { core::ptr::null() } else { self.habitat.as_ptr() },
            alias_ranges: if self.ranges.is_empty() { core::ptr::null() } else { self.ranges.as_ptr() },
            self_dispersal: if self.self_dispersal.is_empty() {
                core::ptr::null()
            } else {
                self.self_dispersal.as_ptr()
            },
            atoms: if self.atoms.is_empty() { core::ptr::null() } else { self.atoms.as_ptr() },
            n_atoms: self.atoms.len(),
            ..self.desc
        }
    }
}

pub trait HipScenario<M: MathsCore>: Scenario<M, WyHash<M>> {
    /// The device description of the cogs, or an error if the device library
    /// does not implement this scenario's cogs.
    ///
    /// # Errors
    ///
    /// Returns `HipError` (NH_E_INVALID) for unsupported scenarios and for
    /// extents the device cannot index.
    fn hip_cogs(cogs: &ScenarioCogs<M, WyHash<M>, Self>) -> Result<HipCogs, HipError>;

    /// The scenario's origin sampler as the device runs it
    /// (nh_job_add_origins): `Some` with the almost-infinite circle sample's
    /// (centre x, centre y, radius), or (0, 0, 0) for the non-spatial and
    /// in-memory samplers; `None` where the device has no such sampler, and
    /// the reference's own walk runs on the host.
    fn device_origins(cogs: &ScenarioCogs<M, WyHash<M>, Self>) -> Option<(u32, u32, u32)> {
        let _ = cogs;
        None
    }
}

fn invalid(message: String) -> HipError {
    HipError { code: ffi::NH_E_INVALID, message }
}

fn to_u32(v: u64, what: &str) -> Result<u32, HipError> {
    u32::try_from(v).map_err(|_| invalid(format!("{what} {v} does not fit the device's 32-bit extent")))
}

/// UniformSpeciationProbability / UniformTurnoverRate evaluated at the
/// extent's origin (both are the same at every location)
fn uniform_rates<M: MathsCore, H: Habitat<M>, T: TurnoverRate<M, H>, N: SpeciationProbability<M, H>>(
    habitat: &H,
    turnover_rate: &T,
    speciation_probability: &N,
) -> (f64, f64) {
    let origin = habitat.get_extent().origin().clone();
    (
        speciation_probability.get_speciation_probability_at_location(&origin, habitat).get(),
        turnover_rate.get_turnover_rate_at_location(&origin, habitat).get(),
    )
}

/// extent (width, height, origin x, origin y) of a bounded habitat
fn extent_of<M: MathsCore, H: Habitat<M>>(habitat: &H) -> Result<(u32, u32, u32, u32), HipError> {
    let e = habitat.get_extent();
    let (w, h) = (to_u32(e.width().get(), "width")?, to_u32(e.height().get(), "height")?);
    if u64::from(w) * u64::from(h) > u64::from(u32::MAX) {
        return Err(invalid(format!("a {w}x{h} habitat exceeds the device's 2^32 cells")));
    }
    Ok((w, h, e.origin().x(), e.origin().y()))
}

/// the habitat map row by row (y-major, like InMemoryHabitat's buffer and the
/// extent's location order, necsim/core/src/landscape/extent.rs:102-280)
fn habitat_map<M: MathsCore, H: Habitat<M>>(habitat: &H, w: u32, h: u32, ox: u32, oy: u32) -> Vec<u32> {
    let mut map = Vec::with_capacity((w as usize) * (h as usize));
    for y in 0..h {
        for x in 0..w {
            map.push(habitat.get_habitat_at_location(&Location::new(ox.wrapping_add(x), oy.wrapping_add(y))));
        }
    }
    map
}

#[cfg(feature = "non-spatial-scenario")]
impl<M: MathsCore> HipScenario<M> for rustcoalescence_scenarios::non_spatial::NonSpatialScenario {
    fn hip_cogs(cogs: &ScenarioCogs<M, WyHash<M>, Self>) -> Result<HipCogs, HipError> {
        let (width, height, origin_x, origin_y) = extent_of(&cogs.habitat)?;
        let origin = cogs.habitat.get_extent().origin().clone();
        let deme = cogs.habitat.get_habitat_at_location(&origin);
        let (nu, lambda) = uniform_rates(&cogs.habitat, &cogs.turnover_rate, &cogs.speciation_probability);
        let desc = ffi::nh_scenario {
            kind: ffi::NH_SCENARIO_NON_SPATIAL,
            width,
            height,
            origin_x,
            origin_y,
            deme,
            ..Default::default()
        };
        Ok(HipCogs::new(desc, nu, lambda))
    }

    fn device_origins(_cogs: &ScenarioCogs<M, WyHash<M>, Self>) -> Option<(u32, u32, u32)> {
        // NonSpatialOriginSampler (origin_sampler/non_spatial.rs:89-112)
        Some((0, 0, 0))
    }
}

#[cfg(feature = "almost-infinite-normal-dispersal-scenario")]
impl<M: MathsCore> HipScenario<M>
    for rustcoalescence_scenarios::almost_infinite::normal::AlmostInfiniteNormalDispersalScenario
{
    fn hip_cogs(cogs: &ScenarioCogs<M, WyHash<M>, Self>) -> Result<HipCogs, HipError> {
        let (nu, lambda) = uniform_rates(&cogs.habitat, &cogs.turnover_rate, &cogs.speciation_probability);
        let desc = ffi::nh_scenario {
            kind: ffi::NH_SCENARIO_ALMOST_INFINITE_NORMAL,
            // AlmostInfiniteNormalDispersalSampler::sigma (added by the patch)
            sigma: cogs.dispersal_sampler.sigma().get(),
            ..Default::default()
        };
        Ok(HipCogs::new(desc, nu, lambda))
    }

    fn device_origins(cogs: &ScenarioCogs<M, WyHash<M>, Self>) -> Option<(u32, u32, u32)> {
        // SingletonDemesCircleOriginSampler (singleton_demes/circle.rs:52-156);
        // the rectangle sample has no device sampler
        match &cogs.origin_sampler_auxiliary.0 {
            rustcoalescence_scenarios::almost_infinite::Sample::Circle { centre, radius } => {
                Some((centre.x(), centre.y(), u32::from(*radius)))
            },
            rustcoalescence_scenarios::almost_infinite::Sample::Rectangle(_) => None,
        }
    }
}

#[cfg(feature = "spatially-explicit-uniform-turnover-scenario")]
impl<M: MathsCore> HipScenario<M>
    for rustcoalescence_scenarios::spatially_explicit::uniform::SpatiallyExplicitUniformTurnoverScenario
{
    fn hip_cogs(cogs: &ScenarioCogs<M, WyHash<M>, Self>) -> Result<HipCogs, HipError> {
        let (width, height, origin_x, origin_y) = extent_of(&cogs.habitat)?;
        let (nu, lambda) = uniform_rates(&cogs.habitat, &cogs.turnover_rate, &cogs.speciation_probability);
        let sampler = &cogs.dispersal_sampler;
        let mut out = HipCogs::new(
            ffi::nh_scenario {
                kind: ffi::NH_SCENARIO_SPATIALLY_EXPLICIT,
                width,
                height,
                origin_x,
                origin_y,
                ..Default::default()
            },
            nu,
            lambda,
        );
        out.habitat = habitat_map(&cogs.habitat, width, height, origin_x, origin_y);
        // AliasSamplerRange { start: usize, end: usize } and
        // SeparableAliasSelfDispersal { self_dispersal: ClosedUnitF64,
        // non_self_dispersal_event: usize } are #[repr(C)] with the layouts of
        // nh_alias_range and nh_self_dispersal on 64-bit targets
        out.ranges = sampler
            .alias_dispersal_ranges()
            .elements_row_major_iter()
            .map(|r| {
                debug_assert_eq!(core::mem::size_of_val(r), core::mem::size_of::<ffi::nh_alias_range>());
                unsafe { core::ptr::read((r as *const _).cast::<ffi::nh_alias_range>()) }
            })
            .collect();
        out.self_dispersal = sampler
            .self_dispersal()
            .elements_row_major_iter()
            .map(|s| {
                debug_assert_eq!(core::mem::size_of_val(s), core::mem::size_of::<ffi::nh_self_dispersal>());
                unsafe { core::ptr::read((s as *const _).cast::<ffi::nh_self_dispersal>()) }
            })
            .collect();
        out.atoms = sampler
            .alias_dispersal_buffer()
            .iter()
            .map(|a| ffi::nh_alias_atom { u: a.u().get(), e: a.e() as u64, k: a.k() as u64 })
            .collect();
        if out.ranges.len() != out.habitat.len() || out.self_dispersal.len() != out.habitat.len() {
            return Err(invalid("the dispersal sampler's tables do not cover the habitat extent".into()));
        }
        Ok(out)
    }

    fn device_origins(_cogs: &ScenarioCogs<M, WyHash<M>, Self>) -> Option<(u32, u32, u32)> {
        // InMemoryOriginSampler (origin_sampler/in_memory.rs:89-119)
        Some((0, 0, 0))
    }
}

/// Scenarios whose cogs the device library does not implement: the
/// algorithm refuses them when the simulation starts.
macro_rules! unsupported_scenario {
    ($feature:literal, $ty:path, $why:literal) => {
        #[cfg(feature = $feature)]
        impl<M: MathsCore> HipScenario<M> for $ty {
            fn hip_cogs(_cogs: &ScenarioCogs<M, WyHash<M>, Self>) -> Result<HipCogs, HipError> {
                Err(invalid(format!(
                    "the HIP algorithm does not support the {} scenario: {}",
                    stringify!($ty),
                    $why
                )))
            }
        }
    };
}

unsupported_scenario!(
    "almost-infinite-clark2dt-dispersal-scenario",
    rustcoalescence_scenarios::almost_infinite::clark2dt::AlmostInfiniteClark2DtDispersalScenario,
    "its Clark 2Dt dispersal kernel is not implemented on the device"
);
unsupported_scenario!(
    "spatially-explicit-turnover-map-scenario",
    rustcoalescence_scenarios::spatially_explicit::map::SpatiallyExplicitTurnoverMapScenario,
    "its per-cell turnover rates are not implemented on the device"
);
unsupported_scenario!(
    "spatially-implicit-scenario",
    rustcoalescence_scenarios::spatially_implicit::SpatiallyImplicitScenario,
    "its metacommunity habitat is not implemented on the device"
);
unsupported_scenario!(
    "wrapping-noise-scenario",
    rustcoalescence_scenarios::wrapping_noise::WrappingNoiseScenario,
    "its noise habitat is not implemented on the device"
);
