//! MI355X (gfx950) scoring backend for `gnomon score`, bound to `libgnomon_hip.so`.
//!
//! NOT BUILT IN THIS IMAGE: there is no cargo or rustc here.  This file is the
//! Rust side of the drop-in, written against the reference crate's types, to
//! be placed at `score/hip_backend.rs` of gnomon with `pub mod hip_backend;` in
//! `score/mod.rs` and the call at `score/pipeline.rs:480` changed from
//! `cuda_backend::try_run_cuda(context)` to `hip_backend::try_run_hip(context)`.
//! The library is built by `score/hip_build.rs`.  The Python mirror of the
//! same control flow, which the GPU tests drive, is
//! `gnomon_amd/pipeline.py::try_run_hip`.
//!
//! What it replaces, function by function:
//! * `try_run_hip`       -> `try_run_cuda`          (score/cuda_backend.rs:604-656)
//! * `run_single_file`   -> `run_single_file_cuda`  (score/cuda_backend.rs:1238-1249)
//! * `run_multi_file`    -> `run_multi_file_cuda`   (score/cuda_backend.rs:1251-1265)
//! * `run_hip_stream`    -> `run_cuda_pipeline` + `process_dense_stream_cuda`
//!                          (score/cuda_backend.rs:1279-1540, 1624-2065),
//!                          reading rows as the producers do
//!                          (score/io.rs:156-321 single file, :325-420 multi file)
//!
//! The library owns all device memory, streams and pinned staging; every
//! buffer passed below is a plain host pointer that must stay valid only for
//! the duration of the call.  Results are `[person][score]` f64 sums with the
//! baseline included and u32 missing counts, as the CUDA backend returns them.

use crate::pipeline_error::PipelineError;
use crate::score::checkpoint::ScoreCheckpointWriter;
use crate::score::complex::{ComplexVariantResolver, resolve_complex_variants};
use crate::score::io::{self, BedSource, ByteRangeSource};
use crate::score::pipeline::PipelineContext;
use crate::score::types::{FilesetBoundary, PipelineKind, PreparationResult};

use std::ffi::CStr;
use std::os::raw::{c_char, c_int, c_void};
use std::panic::{AssertUnwindSafe, catch_unwind};
use std::path::Path;
use std::sync::Arc;
use std::sync::mpsc::sync_channel;
use std::time::{Duration, Instant};

// ---- C ABI of include/gnomon_hip.h ---------------------------------------

#[repr(C)]
struct GnmProblem {
    num_people: u64,
    total_people_in_fam: u64,
    bytes_per_variant: u64,
    num_scores: u64,
    num_variants: u64,
    nnz: u64,
    sparse_weights: *const f32,
    sparse_missing_corrections: *const f32,
    sparse_score_columns: *const u32,
    sparse_row_offsets: *const u64,
    output_idx_to_fam_idx: *const u32,
    precision: i32,
    device: i32,
    reserved: [u64; 4],
}

const GNM_OK: c_int = 0;
const GNM_DECLINED: c_int = 1;
const GNM_PRECISION_AUTO: i32 = 3;

#[link(name = "gnomon_hip")]
unsafe extern "C" {
    fn gnm_last_error() -> *const c_char;
    fn gnm_abi_version() -> c_int;
    fn gnm_size_gate(num_people: u64, num_scores: u64, num_variants: u64) -> c_int;
    fn gnm_device_count() -> c_int;
    fn gnm_create(problem: *const GnmProblem, out_ctx: *mut *mut c_void) -> c_int;
    fn gnm_score_rows(ctx: *mut c_void, rows: *const u8, row_stride: u64, first_variant: u64,
                      num_rows: u64) -> c_int;
    fn gnm_read_accumulators(ctx: *mut c_void, scores: *mut f64, counts: *mut u32,
                             completed_variants: *mut u64) -> c_int;
    fn gnm_seed_accumulators(ctx: *mut c_void, scores: *const f64, counts: *const u32,
                             completed_variants: u64) -> c_int;
    fn gnm_finish(ctx: *mut c_void, scores: *mut f64, counts: *mut u32) -> c_int;
    fn gnm_destroy(ctx: *mut c_void);
}

const EXPECTED_ABI: c_int = 1;

/// Rows handed to the library per call.  The library stages host rows through
/// two pinned slots of at most 256 MiB, so blocks of that size keep both slots
/// busy while the reader thread fills the next block.
const BLOCK_BYTES: u64 = 256 << 20;

/// How often the accumulators are read back for a checkpoint.  Reading them
/// waits for the device and copies N*K*12 bytes, so it is not done per block;
/// `ScoreCheckpointWriter::maybe_save` applies its own cadence on top.
const CHECKPOINT_POLL: Duration = Duration::from_secs(30);

fn last_error() -> String {
    // SAFETY: gnm_last_error returns a thread-local NUL-terminated string (never NULL).
    unsafe { CStr::from_ptr(gnm_last_error()) }.to_string_lossy().into_owned()
}

/// Owns a library context; dropping it releases every device resource in the
/// library's fixed order.
struct CtxGuard(*mut c_void);

impl Drop for CtxGuard {
    fn drop(&mut self) {
        // SAFETY: the pointer came from gnm_create and is destroyed once.
        unsafe { gnm_destroy(self.0) };
    }
}

fn status_to_error(what: &str, status: c_int) -> PipelineError {
    PipelineError::Compute(format!("HIP backend: {what} failed (status {status}): {}", last_error()))
}

// ---- entry point ---------------------------------------------------------

/// Same contract as `try_run_cuda` (score/cuda_backend.rs:604-656):
/// `Ok(None)` when the backend declines (size gate, no device, context that
/// cannot be created, a panic before any result), `Ok(Some(..))` on success,
/// `Err(..)` for a failure after computation started.
pub fn try_run_hip(context: &PipelineContext) -> Result<Option<(Vec<f64>, Vec<u32>)>, PipelineError> {
    let prep = &context.prep_result;
    let (n, k, v) = (prep.num_people_to_score as u64, prep.score_names.len() as u64,
                     prep.num_reconciled_variants as u64);
    // SAFETY: plain value arguments.
    if unsafe { gnm_size_gate(n, k, v) } == 0 {
        eprintln!("> Backend: CPU fallback (problem size below GPU threshold)");
        return Ok(None);
    }
    if unsafe { gnm_abi_version() } != EXPECTED_ABI {
        eprintln!("> Backend: CPU fallback (libgnomon_hip ABI mismatch)");
        return Ok(None);
    }
    if unsafe { gnm_device_count() } <= 0 {
        eprintln!("> Backend: CPU fallback (no gfx950 device detected)");
        return Ok(None);
    }
    // Complex rules re-read rows after scoring.  Sources without a memory map
    // would need the CUDA path's spool; the CPU pipeline handles those.
    if !prep.complex_rules.is_empty() && !all_sources_mapped(prep)? {
        eprintln!("> Backend: CPU fallback (complex variants over unmapped .bed sources)");
        return Ok(None);
    }

    let fam_map: Vec<u32> = prep.output_idx_to_fam_idx.iter().map(|i| i.0).collect();
    let problem = GnmProblem {
        num_people: n,
        total_people_in_fam: prep.total_people_in_fam as u64,
        bytes_per_variant: prep.bytes_per_variant,
        num_scores: k,
        num_variants: v,
        nnz: prep.sparse_weights().len() as u64,
        sparse_weights: prep.sparse_weights().as_ptr(),
        sparse_missing_corrections: prep.sparse_missing_corrections().as_ptr(),
        sparse_score_columns: prep.sparse_score_columns().as_ptr(),
        sparse_row_offsets: prep.sparse_row_offsets().as_ptr(),
        output_idx_to_fam_idx: fam_map.as_ptr(),
        precision: GNM_PRECISION_AUTO,
        device: 0,
        reserved: [0; 4],
    };
    let mut raw: *mut c_void = std::ptr::null_mut();
    // SAFETY: problem points into prep and fam_map, both alive for the call.
    let status = unsafe { gnm_create(&problem, &mut raw) };
    if status == GNM_DECLINED {
        eprintln!("> Backend: CPU fallback (problem size below GPU threshold)");
        return Ok(None);
    }
    if status != GNM_OK {
        eprintln!("> Backend: CPU fallback ({})", last_error());
        return Ok(None);
    }
    eprintln!("> Backend: HIP (gfx950)");
    let ctx = CtxGuard(raw);

    let outcome = catch_unwind(AssertUnwindSafe(|| match &prep.pipeline_kind {
        PipelineKind::SingleFile(path) => run_single_file(context, path, &ctx),
        PipelineKind::MultiFile(boundaries) => run_multi_file(context, boundaries, &ctx),
    }));
    match outcome {
        Ok(Ok(result)) => Ok(Some(result)),
        Ok(Err(e)) => Err(e),
        Err(_) => {
            eprintln!("> Backend: CPU fallback (HIP execution panicked)");
            Ok(None)
        }
    }
}

fn all_sources_mapped(prep: &PreparationResult) -> Result<bool, PipelineError> {
    Ok(match &prep.pipeline_kind {
        PipelineKind::SingleFile(path) => io::open_bed_source(path)?.mmap().is_some(),
        PipelineKind::MultiFile(boundaries) => {
            let mut all = true;
            for b in boundaries {
                all &= io::open_bed_source(&b.bed_path)?.mmap().is_some();
            }
            all
        }
    })
}

fn run_single_file(context: &PipelineContext, bed_path: &Path, ctx: &CtxGuard)
                   -> Result<(Vec<f64>, Vec<u32>), PipelineError> {
    let source = io::open_bed_source(bed_path)?;
    let reader = RowReader::single(source.byte_source());
    let (mut scores, mut counts) = run_hip_stream(context, ctx, reader)?;
    if !context.prep_result.complex_rules.is_empty() {
        let resolver = ComplexVariantResolver::from_single_source(source);
        resolve_complex_variants(&resolver, &context.prep_result, &mut scores, &mut counts)?;
    }
    Ok((scores, counts))
}

fn run_multi_file(context: &PipelineContext, boundaries: &[FilesetBoundary], ctx: &CtxGuard)
                  -> Result<(Vec<f64>, Vec<u32>), PipelineError> {
    let sources: Vec<BedSource> =
        boundaries.iter().map(|b| io::open_bed_source(&b.bed_path)).collect::<Result<_, _>>()?;
    let reader = RowReader::multi(boundaries, &sources);
    let (mut scores, mut counts) = run_hip_stream(context, ctx, reader)?;
    if !context.prep_result.complex_rules.is_empty() {
        let resolver = ComplexVariantResolver::from_multi_sources(sources, boundaries.to_vec())?;
        resolve_complex_variants(&resolver, &context.prep_result, &mut scores, &mut counts)?;
    }
    Ok((scores, counts))
}

// ---- row reading ---------------------------------------------------------

/// Where the packed row of a global BIM row index lives: the fileset whose
/// range holds it and the byte offset `3 + local_row * bpv` inside that
/// fileset's `.bed` (score/io.rs:194, 267 and the boundary walk of :345-360).
struct RowReader {
    sources: Vec<Arc<dyn ByteRangeSource>>,
    /// First global BIM index of each fileset, ascending; [0] for one file.
    starts: Vec<u64>,
}

impl RowReader {
    fn single(source: Arc<dyn ByteRangeSource>) -> Self {
        Self { sources: vec![source], starts: vec![0] }
    }

    fn multi(boundaries: &[FilesetBoundary], sources: &[BedSource]) -> Self {
        Self {
            sources: sources.iter().map(|s| s.byte_source()).collect(),
            starts: boundaries.iter().map(|b| b.starting_global_index).collect(),
        }
    }

    fn fileset_of(&self, global_row: u64) -> usize {
        self.starts.partition_point(|&s| s <= global_row).saturating_sub(1)
    }

    /// Fills `dst` with the rows of reconciled variants [first, first + n):
    /// runs of consecutive BIM rows inside one fileset become one read.
    fn fill(&self, prep: &PreparationResult, first: usize, n: usize, dst: &mut [u8])
            -> Result<(), PipelineError> {
        let bpv = prep.bytes_per_variant;
        let bim = &prep.required_bim_indices[first..first + n];
        let mut i = 0;
        while i < n {
            let row = bim[i].0;
            let f = self.fileset_of(row);
            let fileset_end = self.starts.get(f + 1).copied().unwrap_or(u64::MAX);
            let mut run = 1;
            while i + run < n && bim[i + run].0 == row + run as u64 && bim[i + run].0 < fileset_end {
                run += 1;
            }
            let offset = 3 + (row - self.starts[f]) * bpv;
            let len = run as u64 * bpv;
            let source = &self.sources[f];
            if offset + len > source.len() {
                return Err(PipelineError::Io(format!(
                    "Fatal: read past the end of a .bed source for the variant at BIM row {row}. \
                     The file may be truncated or inconsistent with the .bim file."
                )));
            }
            let at = i * bpv as usize;
            source.read_at(offset, &mut dst[at..at + len as usize])?;
            i += run;
        }
        Ok(())
    }
}

// ---- the stream ----------------------------------------------------------

/// Resume, blocks, checkpoints and the final read-back.  A reader thread fills
/// one host block while the library copies the previous one into its pinned
/// staging; `gnm_score_rows` returns as soon as a block's bytes have been
/// consumed, so two blocks in rotation keep the file, PCIe and the device busy.
fn run_hip_stream(context: &PipelineContext, ctx: &CtxGuard, reader: RowReader)
                  -> Result<(Vec<f64>, Vec<u32>), PipelineError> {
    let prep: &PreparationResult = &context.prep_result;
    let total = prep.num_reconciled_variants;
    let cells = prep.num_people_to_score * prep.score_names.len();
    let bpv = prep.bytes_per_variant;
    let mut scores = vec![0.0f64; cells];
    let mut counts = vec![0u32; cells];

    // Resume (score/pipeline.rs:1777-1799, score/cuda_backend.rs:1680-1694).
    let mut done = context.checkpoint_completed_variants();
    if let Some(cp) = context.checkpoint.as_ref() {
        if cp.sum_scores.len() != cells || cp.missing_counts.len() != cells {
            return Err(PipelineError::Compute(format!(
                "Checkpoint holds {} scores and {} counts; this problem has {cells} cells.",
                cp.sum_scores.len(), cp.missing_counts.len())));
        }
        // SAFETY: both slices hold `cells` values.
        let st = unsafe { gnm_seed_accumulators(ctx.0, cp.sum_scores.as_ptr(),
                                                cp.missing_counts.as_ptr(), done as u64) };
        if st != GNM_OK {
            return Err(status_to_error("gnm_seed_accumulators", st));
        }
        eprintln!("> Resuming HIP score computation from checkpoint at {done}/{total} variants.");
    }
    let mut writer = match (context.checkpoint_path.as_ref(), context.checkpoint_fingerprint) {
        (Some(path), Some(fp)) => Some(ScoreCheckpointWriter::new(path.clone(), fp, total)),
        _ => None,
    };

    let rows_per_block = ((BLOCK_BYTES / bpv.max(1)).max(1) as usize).min(total.max(1));
    let (full_tx, full_rx) = sync_channel::<Result<(usize, usize, Vec<u8>), PipelineError>>(1);
    let (empty_tx, empty_rx) = sync_channel::<Vec<u8>>(2);
    for _ in 0..2 {
        empty_tx.send(vec![0u8; rows_per_block * bpv as usize]).expect("fresh channel");
    }

    std::thread::scope(|scope| -> Result<(), PipelineError> {
        let reader = &reader;
        let start = done;
        scope.spawn(move || {
            let mut first = start;
            while first < total {
                let Ok(mut buf) = empty_rx.recv() else { return };
                let n = rows_per_block.min(total - first);
                let filled = reader.fill(prep, first, n, &mut buf).map(|()| (first, n, buf));
                let failed = filled.is_err();
                if full_tx.send(filled).is_err() || failed {
                    return;
                }
                first += n;
            }
        });

        let mut last_poll = Instant::now();
        while done < total {
            let (first, n, buf) = full_rx.recv().map_err(|_| {
                PipelineError::Compute("HIP backend: the row reader stopped early.".to_string())
            })??;
            debug_assert_eq!(first, done);
            // SAFETY: buf holds n rows of bpv bytes; the call returns once they are consumed.
            let st = unsafe { gnm_score_rows(ctx.0, buf.as_ptr(), bpv, first as u64, n as u64) };
            let _ = empty_tx.send(buf);
            if st != GNM_OK {
                return Err(status_to_error("gnm_score_rows", st));
            }
            done += n;
            if let Some(w) = writer.as_mut() {
                if last_poll.elapsed() >= CHECKPOINT_POLL && done < total {
                    let mut completed = 0u64;
                    // SAFETY: both outputs hold `cells` values.
                    let st = unsafe { gnm_read_accumulators(ctx.0, scores.as_mut_ptr(),
                                                            counts.as_mut_ptr(), &mut completed) };
                    if st != GNM_OK {
                        return Err(status_to_error("gnm_read_accumulators", st));
                    }
                    save_checkpoint(w, completed as usize, &scores, &counts, false)?;
                    last_poll = Instant::now();
                }
            }
        }
        drop(empty_tx);
        Ok(())
    })?;

    // SAFETY: both outputs hold `cells` values.
    let st = unsafe { gnm_finish(ctx.0, scores.as_mut_ptr(), counts.as_mut_ptr()) };
    if st != GNM_OK {
        return Err(status_to_error("gnm_finish", st));
    }
    if let Some(w) = writer.as_mut() {
        save_checkpoint(w, total, &scores, &counts, true)?;
    }
    Ok((scores, counts))
}

fn save_checkpoint(writer: &mut ScoreCheckpointWriter, completed: usize, scores: &[f64], counts: &[u32],
                   force: bool) -> Result<(), PipelineError> {
    let wrote = writer
        .maybe_save(completed, scores, counts, force)
        .map_err(|e| PipelineError::Io(format!("Failed to write score checkpoint: {e}")))?;
    if wrote {
        eprintln!("> Score checkpoint saved after {completed} completed variants.");
    }
    Ok(())
}
