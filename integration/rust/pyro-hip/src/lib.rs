//! Rust binding of the MI355X fire solver (include/pyro_hip.h).
//!
//! `Sim` keeps the reference's method names and meaning (src/sim.rs:108
//! `Sim::new`, :246 `step`, :494 `read_scalars`, :510 `read_stats`, :234
//! `cells`) without the `&Gpu` argument: the library owns the device, its
//! memory and its stream.  The caller passes the 576-byte `Params` blob
//! (scene.rs:423-477, `repr(C)`) and the voxelized SOL (scene.rs:621) that
//! the reference's `Scene` already produces.  A negative return code of the
//! C ABI becomes an `anyhow` error carrying `pyro_last_error()`.
use std::ffi::CStr;
use std::os::raw::{c_char, c_int, c_void};

#[repr(C)]
pub struct PyroSim {
    _p: [u8; 0],
}

pub const PARAMS_BYTES: usize = 576;
pub const STATS_N: usize = 1040;
/// Field ids of pyro_sim_read_field (include/pyro_hip.h).
pub const FIELD_S0: c_int = 0;
pub const FIELD_S1: c_int = 1;

extern "C" {
    fn pyro_last_error() -> *const c_char;
    fn pyro_sim_create(params: *const c_void, params_bytes: usize, sol_init: *const f32, s0_init: *const f32,
                       pressure_cycles: u32, ign_until: f32, flags: u32, out: *mut *mut PyroSim) -> c_int;
    fn pyro_sim_destroy(sim: *mut PyroSim) -> c_int;
    fn pyro_sim_step(sim: *mut PyroSim, dt: f32, want_stats: c_int) -> c_int;
    fn pyro_sim_read_stats(sim: *mut PyroSim, out: *mut f64, n: usize) -> c_int;
    fn pyro_sim_read_field(sim: *mut PyroSim, field: c_int, dst: *mut c_void, bytes: usize) -> c_int;
    fn pyro_sim_sync(sim: *mut PyroSim) -> c_int;
}

fn check(rc: c_int) -> anyhow::Result<()> {
    if rc == 0 {
        return Ok(());
    }
    let msg = unsafe { CStr::from_ptr(pyro_last_error()) }.to_string_lossy().into_owned();
    Err(anyhow::anyhow!("pyro_hip error {rc}: {msg}"))
}

pub type Field4 = Vec<[f32; 4]>;

/// Statistics of a step with `want_stats` (stats.wgsl:2-3 slots, FP64 sums
/// narrowed to the reference's f32 `Stats`, sim.rs:73-87).
#[derive(Clone, Debug, Default)]
pub struct Stats {
    pub hrr: f32,
    pub rad: f32,
    pub pyro: f32,
    pub soot: f32,
    pub max_t: f32,
    pub max_u: f32,
    pub fuel: f32,
    pub solid: f32,
    pub smoke: f32,
    pub area: f32,
    pub resid: f32,
    pub layers: Vec<f32>,
}

pub struct Sim {
    h: *mut PyroSim,
    pub dims: [u32; 3],
    pub dx: f32,
    pub time: f32,
    pub cycles: u32,
}

impl Sim {
    /// sim.rs:108.  `params` is the scene's `Params` as bytes, `sol` its
    /// voxelization (nx*ny*nz cells of 4 f32).
    pub fn new(params: &[u8], sol: &[[f32; 4]], pressure_cycles: u32, ign_until: f32) -> anyhow::Result<Sim> {
        anyhow::ensure!(params.len() == PARAMS_BYTES, "Params must be {PARAMS_BYTES} bytes");
        let u = |o: usize| u32::from_le_bytes([params[o], params[o + 1], params[o + 2], params[o + 3]]);
        let dims = [u(0), u(4), u(8)];
        let dx = f32::from_bits(u(16));
        let n = dims.iter().map(|&d| d as usize).product::<usize>();
        anyhow::ensure!(sol.len() == n, "SOL holds {} cells, the grid {n}", sol.len());
        let mut h = std::ptr::null_mut();
        check(unsafe {
            pyro_sim_create(params.as_ptr() as *const c_void, PARAMS_BYTES, sol.as_ptr() as *const f32,
                            std::ptr::null(), pressure_cycles, ign_until, 0, &mut h)
        })?;
        Ok(Sim { h, dims, dx, time: 0.0, cycles: pressure_cycles })
    }

    /// sim.rs:246: enqueues one step and returns; a device error surfaces at
    /// the next blocking call, as wgpu's do.
    pub fn step(&mut self, dt: f32, want_stats: bool) {
        let _ = unsafe { pyro_sim_step(self.h, dt, want_stats as c_int) };
        self.time += dt;
    }

    /// sim.rs:510: blocks; valid after a step with `want_stats`.
    pub fn read_stats(&self) -> anyhow::Result<Stats> {
        let mut v = vec![0f64; STATS_N];
        check(unsafe { pyro_sim_read_stats(self.h, v.as_mut_ptr(), STATS_N) })?;
        let ny = (self.dims[1] as usize).min(1024);
        Ok(Stats {
            hrr: v[0] as f32, rad: v[1] as f32, pyro: v[2] as f32, soot: v[3] as f32, max_t: v[4] as f32,
            max_u: v[5] as f32, fuel: v[6] as f32, solid: v[7] as f32, smoke: v[8] as f32, area: v[9] as f32,
            resid: if v[11] > 0.0 { (v[10] / v[11]).sqrt() as f32 } else { 0.0 },
            layers: v[16..16 + ny].iter().map(|x| *x as f32).collect(),
        })
    }

    /// sim.rs:494: (s0, s1), blocks.
    pub fn read_scalars(&self) -> anyhow::Result<(Field4, Field4)> {
        let n = self.cells() as usize;
        let (mut a, mut b) = (vec![[0f32; 4]; n], vec![[0f32; 4]; n]);
        check(unsafe { pyro_sim_read_field(self.h, FIELD_S0, a.as_mut_ptr() as *mut c_void, n * 16) })?;
        check(unsafe { pyro_sim_read_field(self.h, FIELD_S1, b.as_mut_ptr() as *mut c_void, n * 16) })?;
        Ok((a, b))
    }

    /// sim.rs:234
    pub fn cells(&self) -> u64 {
        self.dims.iter().map(|&d| d as u64).product()
    }

    /// Waits for every queued step.
    pub fn sync(&self) -> anyhow::Result<()> {
        check(unsafe { pyro_sim_sync(self.h) })
    }
}

impl Drop for Sim {
    fn drop(&mut self) {
        unsafe {
            pyro_sim_destroy(self.h);
        }
    }
}
