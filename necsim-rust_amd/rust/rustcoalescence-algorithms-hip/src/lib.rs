//! `Independent`-semantics algorithm on an AMD MI355X through libnecsim_hip.so.
//!
//! Mirrors rustcoalescence/algorithms/independent/src/lib.rs:39-173 (CPU
//! Independent) and rustcoalescence/algorithms/cuda/src/lib.rs:66-593 (the
//! existing GPU drop-in): the CPU algorithm's arguments (delta_t, step_slice,
//! dedup_cache, Monolithic event_slice), the deterministic MathsCore
//! (ReproducibleMathsCore) and RNG (WyHash), the same `SimulationOutcome`.
//! Every step runs on the GPU; each water-level round's events come back
//! sorted and are reported through the local partition like the live
//! water-level reporter
//! (necsim/impls/no-std/src/parallelisation/independent/monolithic/reporter/live.rs:88-115),
//! and progress (the lineages left) is reported like the monolithic loop does
//! (.../parallelisation/independent/monolithic/mod.rs:107,244,260), once per
//! round instead of once per pass.

use std::{ffi::CStr, marker::PhantomData};

use necsim_core::{
    cogs::MathsCore,
    event::{PackedEvent, SpeciationEvent, TypedEvent},
    lineage::Lineage,
    reporter::{boolean::Boolean, Reporter},
};
use necsim_core_bond::{NonNegativeF64, PositiveF64};
use necsim_impls_no_std::cogs::{
    active_lineage_sampler::resuming::lineage::ExceptionalLineage,
    lineage_store::independent::IndependentLineageStore, maths::reproducible::ReproducibleMathsCore,
    origin_sampler::{pre_sampler::OriginPreSampler, resuming::ResumingOriginSampler},
    rng::wyhash::WyHash,
};
use necsim_partitioning_core::{
    partition::{Partition, PartitionSize},
    LocalPartition, Partitioning,
};
use rustcoalescence_algorithms::{
    result::{ResumeError, SimulationOutcome},
    strategy::RestartFixUpStrategy,
    Algorithm, AlgorithmDefaults, AlgorithmDispatch, AlgorithmParamters,
};
use rustcoalescence_scenarios::ScenarioCogs;

mod arguments;
pub mod ffi;
mod scenario;

pub use arguments::{HipArguments, HipForm};
pub use scenario::{HipCogs, HipScenario};

#[derive(thiserror::Error, Debug, serde::Serialize, serde::Deserialize)]
#[error("necsim-hip error {code}: {message}")]
pub struct HipError {
    pub code: i32,
    pub message: String,
}

fn check(code: i32) -> Result<(), HipError> {
    if code == ffi::NH_OK {
        return Ok(());
    }
    let mut buf = [0 as core::ffi::c_char; 512];
    unsafe { ffi::nh_last_error(buf.as_mut_ptr(), buf.len()) };
    let message = unsafe { CStr::from_ptr(buf.as_ptr()) }.to_string_lossy().into_owned();
    Err(HipError { code, message })
}

#[allow(clippy::module_name_repetitions, clippy::empty_enum)]
pub enum HipAlgorithm {}

impl AlgorithmParamters for HipAlgorithm {
    type Arguments = HipArguments;
    type Error = HipError;
}

impl AlgorithmDefaults for HipAlgorithm {
    // the deterministic libm core; the device restates the same functions
    type MathsCore = ReproducibleMathsCore;
    type Rng<M: MathsCore> = WyHash<M>;
}

impl<M: MathsCore, O: HipScenario<M>, R: Reporter> AlgorithmDispatch<M, WyHash<M>, O, R> for HipAlgorithm {
    type Algorithm<'p, P: LocalPartition<'p, R>> = Self;

    fn get_logical_partition_size<P: Partitioning>(_args: &Self::Arguments, _partitioning: &P) -> PartitionSize {
        PartitionSize::MONOLITHIC
    }
}

/// The WyHash seed (WyHash is #[repr(C)] { seed: u64, state: u64 },
/// necsim/impls/no-std/src/cogs/rng/wyhash.rs:15-24; the state is re-primed
/// before every use, so the seed is all the device needs).
fn wyhash_seed<M: MathsCore>(rng: &WyHash<M>) -> u64 {
    unsafe { *(rng as *const WyHash<M>).cast::<u64>() }
}

/// Lineage and nh_lineage share one #[repr(C)] layout
/// (necsim/core/src/lineage.rs:97-109).
fn as_nh_lineages(lineages: &[Lineage]) -> *const ffi::nh_lineage {
    debug_assert_eq!(core::mem::size_of::<Lineage>(), core::mem::size_of::<ffi::nh_lineage>());
    lineages.as_ptr().cast()
}

/// One device context with the scenario's resident tables (and the job).
struct Device {
    ctx: *mut ffi::nh_ctx,
    tables: *mut ffi::nh_tables,
    job: *mut ffi::nh_job,
}

impl Device {
    fn open(args: &HipArguments, cogs: &HipCogs) -> Result<Self, HipError> {
        let mut dev = Device { ctx: core::ptr::null_mut(), tables: core::ptr::null_mut(), job: core::ptr::null_mut() };
        check(unsafe { ffi::nh_init(args.device, &mut dev.ctx) })?;
        let desc = cogs.scenario();
        check(unsafe { ffi::nh_tables_create(dev.ctx, &desc, &mut dev.tables) })?;
        Ok(dev)
    }

    /// IndependentActiveLineageSampler::resume_with_store_and_lineages's
    /// filters (active_lineage_sampler/independent/mod.rs:98-117): the
    /// lineages that are out of their habitat or deme, in input order.
    fn exceptional(&self, lineages: &[Lineage]) -> Result<Vec<ExceptionalLineage>, HipError> {
        let mut kinds = vec![ffi::NH_LINEAGE_OK; lineages.len()];
        let mut n_bad = 0_usize;
        check(unsafe {
            ffi::nh_classify_lineages(
                self.tables,
                as_nh_lineages(lineages),
                lineages.len(),
                kinds.as_mut_ptr(),
                &mut n_bad,
            )
        })?;
        let mut out = Vec::with_capacity(n_bad);
        for (lineage, kind) in lineages.iter().zip(kinds) {
            match kind {
                ffi::NH_LINEAGE_OUT_OF_DEME => out.push(ExceptionalLineage::OutOfDeme(lineage.clone())),
                ffi::NH_LINEAGE_OUT_OF_HABITAT => out.push(ExceptionalLineage::OutOfHabitat(lineage.clone())),
                _ => (),
            }
        }
        Ok(out)
    }
}

impl Drop for Device {
    fn drop(&mut self) {
        unsafe {
            if !self.job.is_null() {
                ffi::nh_job_destroy(self.job);
            }
            if !self.tables.is_null() {
                ffi::nh_tables_destroy(self.tables);
            }
            if !self.ctx.is_null() {
                ffi::nh_ctx_destroy(self.ctx);
            }
        }
    }
}

/// The autotuner's launch shape for the persistent form (nh_autotune on
/// every 8th lineage over `window`), kept only for the knobs it tunes.
fn autotuned(dev: &Device, params: &ffi::nh_job_params, lineages: &[Lineage], window: f64) -> Result<ffi::nh_job_params, HipError> {
    let sample: Vec<Lineage> = lineages.iter().step_by(8).cloned().collect();
    let mut tuned = *params;
    let mut best_ms = 0.0_f64;
    check(unsafe {
        ffi::nh_autotune(dev.ctx, dev.tables, params, as_nh_lineages(&sample), sample.len(), window, &mut tuned, &mut best_ms)
    })?;
    Ok(tuned)
}

/// Where a job's lineages come from.
enum Origins<'a> {
    /// lineages on the host: resumed ones, or a sampler the device does not run
    Host(&'a [Lineage]),
    /// the scenario's origin sampler run on the device into the job
    /// (nh_job_add_origins), with the pre-sampler's percentage and the circle
    Device { percentage: f64, circle: (u32, u32, u32) },
}

#[allow(clippy::too_many_arguments)]
fn simulate<'p, M: MathsCore, R: Reporter, P: LocalPartition<'p, R>>(
    args: &HipArguments,
    rng: WyHash<M>,
    cogs: &HipCogs,
    mut dev: Device,
    origins: Origins<'_>,
    resume_after: Option<NonNegativeF64>,
    pause_before: Option<NonNegativeF64>,
    local_partition: &mut P,
) -> Result<SimulationOutcome<M, WyHash<M>>, HipError> {
    let mut params = args.job_params(
        wyhash_seed(&rng),
        cogs.speciation_probability,
        cogs.turnover_rate,
        resume_after,
        pause_before,
    );
    // the reporter's ReportDispersal selects the reported event rate and
    // stream (monolithic/mod.rs:139-164)
    let report_dispersal = R::ReportDispersal::VALUE;
    params.report_dispersal = i32::from(report_dispersal);
    if let Origins::Host(lineages) = origins {
        if args.form == HipForm::Persistent && args.autotune_window > 0.0 && !lineages.is_empty() {
            params = autotuned(&dev, &params, lineages, args.autotune_window)?;
        }
    }
    check(unsafe { ffi::nh_job_create(dev.ctx, dev.tables, &params, &mut dev.job) })?;
    let n_initial = match origins {
        Origins::Host(lineages) => {
            check(unsafe { ffi::nh_job_add_lineages(dev.job, as_nh_lineages(lineages), lineages.len()) })?;
            lineages.len() as u64
        },
        Origins::Device { percentage, circle: (cx, cy, radius) } => {
            // the CLI's OriginPreSampler::all().percentage(p), unpartitioned
            // (rustcoalescence/src/cli/simulate/dispatch/valid/launch.rs:34-45)
            check(unsafe { ffi::nh_job_add_origins(dev.job, percentage, 0, 1, cx, cy, radius) })?;
            let mut o = ffi::nh_outcome::default();
            check(unsafe { ffi::nh_job_outcome(dev.job, &mut o) })?;
            o.n_remaining
        },
    };

    // the progress bar starts at the lineages to simulate (monolithic/mod.rs:107)
    local_partition.report_progress_sync(n_initial);

    let mut buf: Vec<ffi::nh_speciation_event> = Vec::new();
    let mut packed: Vec<ffi::nh_packed_event> = Vec::new();
    let mut o = ffi::nh_outcome::default();
    loop {
        let mut finished = 0_i32;
        check(unsafe { ffi::nh_job_step_round(dev.job, &mut finished) })?;
        let mut n = 0_usize;
        if report_dispersal {
            // every event of the round, sorted; dispatched like the live
            // water-level reporter (monolithic/reporter/live.rs:88-106)
            check(unsafe { ffi::nh_job_drain_packed_events(dev.job, core::ptr::null_mut(), 0, &mut n) })?;
            packed.resize(n, ffi::nh_packed_event::default());
            check(unsafe { ffi::nh_job_drain_packed_events(dev.job, packed.as_mut_ptr(), n, &mut n) })?;
            for e in &packed[..n] {
                // PackedEvent is #[repr(C)] with the nh_packed_event layout (event.rs:16-29)
                let event: PackedEvent = unsafe { core::ptr::read((e as *const ffi::nh_packed_event).cast()) };
                match event.into() {
                    TypedEvent::Speciation(event) => local_partition.get_reporter().report_speciation(&event.into()),
                    TypedEvent::Dispersal(event) => local_partition.get_reporter().report_dispersal(&event.into()),
                }
            }
        } else {
            check(unsafe { ffi::nh_job_drain_events(dev.job, core::ptr::null_mut(), 0, &mut n) })?;
            buf.resize(n, ffi::nh_speciation_event::default());
            check(unsafe { ffi::nh_job_drain_events(dev.job, buf.as_mut_ptr(), n, &mut n) })?;
            for e in &buf[..n] {
                // SpeciationEvent is #[repr(C)] with the nh_speciation_event layout (event.rs:58-71)
                let event: &SpeciationEvent = unsafe { &*(e as *const ffi::nh_speciation_event).cast() };
                local_partition.get_reporter().report_speciation(event.into());
            }
        }
        check(unsafe { ffi::nh_job_outcome(dev.job, &mut o) })?;
        // the lineages left after the round (monolithic/mod.rs:244-250)
        local_partition.get_reporter().report_progress(&o.n_remaining.into());
        if finished != 0 {
            break;
        }
    }
    local_partition.report_progress_sync(o.n_remaining);

    let time = NonNegativeF64::new(o.time).unwrap_or(NonNegativeF64::zero());
    if o.paused == 0 {
        return Ok(SimulationOutcome::Done { time, steps: o.steps });
    }
    let mut rem = vec![ffi::nh_lineage::default(); o.n_remaining as usize];
    let mut n = 0_usize;
    check(unsafe { ffi::nh_job_take_lineages(dev.job, rem.as_mut_ptr(), rem.len(), &mut n) })?;
    let lineages = rem[..n]
        .iter()
        .map(|l| unsafe { core::ptr::read((l as *const ffi::nh_lineage).cast::<Lineage>()) })
        .collect();
    Ok(SimulationOutcome::Paused { time, steps: o.steps, lineages, rng, marker: PhantomData::<M> })
}

impl<'p, M: MathsCore, O: HipScenario<M>, R: Reporter, P: LocalPartition<'p, R>> Algorithm<'p, M, WyHash<M>, O, R, P>
    for HipAlgorithm
{
    type LineageStore = IndependentLineageStore<M, O::Habitat>;

    fn get_logical_partition(_args: &Self::Arguments, _local_partition: &P) -> Partition {
        // the monolithic water-level loop (the CPU algorithm's Monolithic mode)
        Partition::monolithic()
    }

    fn initialise_and_simulate<I: Iterator<Item = u64>>(
        args: Self::Arguments,
        rng: WyHash<M>,
        scenario: ScenarioCogs<M, WyHash<M>, O>,
        pre_sampler: OriginPreSampler<M, I>,
        pause_before: Option<NonNegativeF64>,
        local_partition: &mut P,
    ) -> Result<SimulationOutcome<M, WyHash<M>>, Self::Error> {
        let cogs = O::hip_cogs(&scenario)?;
        let dev = Device::open(&args, &cogs)?;
        // the genesis origin walk (independent/src/launch.rs:78-86) runs on the
        // device where it has the scenario's sampler; the autotuner, which
        // times a sample of the lineages, and samplers the device does not
        // run take the reference's walk on the host
        match O::device_origins(&scenario) {
            Some(circle) if args.autotune_window <= 0.0 => {
                let percentage = pre_sampler.get_sample_proportion().get();
                simulate(&args, rng, &cogs, dev, Origins::Device { percentage, circle }, None, pause_before,
                         local_partition)
            },
            _ => {
                let lineages: Vec<Lineage> =
                    O::sample_habitat(&scenario.habitat, pre_sampler, scenario.origin_sampler_auxiliary.clone())
                        .collect();
                simulate(&args, rng, &cogs, dev, Origins::Host(&lineages), None, pause_before, local_partition)
            },
        }
    }

    fn resume_and_simulate<I: Iterator<Item = u64>, L: ExactSizeIterator<Item = Lineage>>(
        args: Self::Arguments,
        rng: WyHash<M>,
        scenario: ScenarioCogs<M, WyHash<M>, O>,
        pre_sampler: OriginPreSampler<M, I>,
        lineages: L,
        resume_after: Option<NonNegativeF64>,
        pause_before: Option<NonNegativeF64>,
        local_partition: &mut P,
    ) -> Result<SimulationOutcome<M, WyHash<M>>, ResumeError<Self::Error>> {
        let cogs = O::hip_cogs(&scenario)?;
        // the pre-sampler picks this run's share of the resumed lineages
        // (independent/src/initialiser/resume.rs:71-74, origin_sampler/resuming.rs:107-120)
        let lineages: Vec<Lineage> = ResumingOriginSampler::new(&scenario.habitat, pre_sampler, lineages).collect();
        let dev = Device::open(&args, &cogs)?;
        // lineages out of their habitat or deme fail the resume
        // (independent/src/initialiser/resume.rs:83-85)
        let exceptional = dev.exceptional(&lineages)?;
        if !exceptional.is_empty() {
            return Err(ResumeError::Sample(exceptional));
        }
        Ok(simulate(&args, rng, &cogs, dev, Origins::Host(&lineages), resume_after, pause_before, local_partition)?)
    }

    fn fixup_for_restart<I: Iterator<Item = u64>, L: ExactSizeIterator<Item = Lineage>>(
        _args: Self::Arguments,
        _rng: WyHash<M>,
        _scenario: ScenarioCogs<M, WyHash<M>, O>,
        _pre_sampler: OriginPreSampler<M, I>,
        _lineages: L,
        _restart_at: PositiveF64,
        _fixup_strategy: RestartFixUpStrategy,
        _local_partition: &mut P,
    ) -> Result<SimulationOutcome<M, WyHash<M>>, ResumeError<Self::Error>> {
        // the CPU Independent algorithm's FixUp path (const event-time sampler,
        // trespassing dispersal) is not part of the device library
        Err(ResumeError::Simulate(HipError {
            code: ffi::NH_E_STATE,
            message: "fixup_for_restart is not offloaded to the device".into(),
        }))
    }
}
