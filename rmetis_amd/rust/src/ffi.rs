//! extern "C" declarations of librmetis_hip.so (include/rmetis.h).
#![allow(non_snake_case)]
use std::os::raw::{c_char, c_void};

pub type Idx = i32;
pub type Real = f32;

pub const METIS_OK: i32 = 1;
pub const METIS_ERROR_INPUT: i32 = -2;
pub const METIS_ERROR_MEMORY: i32 = -3;
pub const METIS_ERROR: i32 = -4;
pub const METIS_NOPTIONS: usize = 40;

#[link(name = "rmetis_hip")]
extern "C" {
    pub fn RMETIS_SetDefaultOptions(options: *mut Idx) -> i32;
    pub fn RMETIS_PartGraphKway(
        nvtxs: *const Idx, ncon: *const Idx, xadj: *const Idx, adjncy: *const Idx, vwgt: *const Idx,
        vsize: *const Idx, adjwgt: *const Idx, nparts: *const Idx, tpwgts: *const Real, ubvec: *const Real,
        options: *const Idx, objval: *mut Idx, part: *mut Idx,
    ) -> i32;
    pub fn RMETIS_PartGraphRecursive(
        nvtxs: *const Idx, ncon: *const Idx, xadj: *const Idx, adjncy: *const Idx, vwgt: *const Idx,
        vsize: *const Idx, adjwgt: *const Idx, nparts: *const Idx, tpwgts: *const Real, ubvec: *const Real,
        options: *const Idx, objval: *mut Idx, part: *mut Idx,
    ) -> i32;
    pub fn RMETIS_NodeND(
        nvtxs: *const Idx, xadj: *const Idx, adjncy: *const Idx, vwgt: *const Idx, options: *const Idx,
        perm: *mut Idx, iperm: *mut Idx,
    ) -> i32;
    pub fn rmetis_hip_validate_csr(
        nvtxs: Idx, ncon: Idx, xadj: *const Idx, xadj_len: i64, adjncy: *const Idx, adjncy_len: i64,
        vwgt: *const Idx, vwgt_len: i64, adjwgt: *const Idx, adjwgt_len: i64,
    ) -> i32;
    pub fn rmetis_hip_edge_cut(
        nvtxs: Idx, xadj: *const Idx, adjncy: *const Idx, adjwgt: *const Idx, part: *const Idx, cut: *mut Idx,
    ) -> i32;
    pub fn rmetis_hip_partition_weights(
        nvtxs: Idx, ncon: Idx, vwgt: *const Idx, part: *const Idx, nparts: Idx, pw: *mut Idx,
    ) -> i32;
    pub fn rmetis_read_graph(
        path: *const c_char, nvtxs: *mut Idx, ncon: *mut Idx, xadj: *mut *mut Idx, adjncy: *mut *mut Idx,
        vwgt: *mut *mut Idx, vsize: *mut *mut Idx, adjwgt: *mut *mut Idx,
    ) -> i32;
    pub fn rmetis_write_graph(
        path: *const c_char, nvtxs: Idx, ncon: Idx, xadj: *const Idx, adjncy: *const Idx, vwgt: *const Idx,
        vsize: *const Idx, adjwgt: *const Idx,
    ) -> i32;
    pub fn rmetis_free(p: *mut c_void);
    pub fn rmetis_hip_last_error() -> *const c_char;
    pub fn rmetis_hip_device_count() -> i32;
}
