"""ctypes signatures of the C-ABI in include/ofeli_hip.h."""
import ctypes as C

vp = C.c_void_p
i32 = C.c_int
i64 = C.c_int64
f64 = C.c_double
i32p = C.POINTER(C.c_int32)
i64p = C.POINTER(C.c_int64)
f64p = C.POINTER(C.c_double)
f32p = C.POINTER(C.c_float)

# name -> (restype, argtypes); keep in the order of include/ofeli_hip.h
SIGNATURES = {
    "ofh_last_error": (C.c_char_p, []),
    "ofh_ctx_create": (i32, [i32, C.POINTER(vp)]),
    "ofh_ctx_destroy": (i32, [vp]),
    "ofh_ctx_set_stream": (i32, [vp, vp]),
    "ofh_ctx_synchronize": (i32, [vp]),
    "ofh_ctx_use_graphs": (i32, [vp, i32]),
    "ofh_ctx_mixed_form": (i32, [vp, i32]),
    "ofh_malloc": (i32, [vp, C.c_size_t, C.POINTER(vp)]),
    "ofh_free": (i32, [vp, vp]),
    "ofh_memcpy_h2d": (i32, [vp, vp, vp, C.c_size_t]),
    "ofh_memcpy_d2h": (i32, [vp, vp, vp, C.c_size_t]),
    "ofh_memcpy_d2d": (i32, [vp, vp, vp, C.c_size_t]),
    "ofh_csr_create": (i32, [vp, i32, i64, i32p, i32p, f64p, C.POINTER(vp)]),
    "ofh_csr_destroy": (i32, [vp]),
    "ofh_csr_info": (i32, [vp, C.POINTER(i32), C.POINTER(i64), C.POINTER(i32)]),
    "ofh_csr_get_pattern": (i32, [vp, i32p, i32p]),
    "ofh_csr_set_values": (i32, [vp, f64p]),
    "ofh_csr_set_values_device": (i32, [vp, vp]),
    "ofh_csr_get_values": (i32, [vp, f64p]),
    "ofh_csr_values_device": (i32, [vp, C.POINTER(vp)]),
    "ofh_csr_compress_columns": (i32, [vp, i32]),
    "ofh_csr_sell_info": (i32, [vp, C.POINTER(C.c_int), C.POINTER(C.c_int64), C.POINTER(C.c_int)]),
    "ofh_csr_dia_info": (i32, [vp, C.POINTER(C.c_int), C.POINTER(C.c_int), C.POINTER(C.c_int),
                               C.POINTER(C.c_int), C.POINTER(C.c_int)]),
    "ofh_spmv_f64": (i32, [vp, vp, vp]),
    "ofh_spmv_add_f64": (i32, [vp, f64, vp, vp]),
    "ofh_spmv_t_f64": (i32, [vp, vp, vp]),
    "ofh_spmv_f32": (i32, [vp, vp, vp]),
    "ofh_dot_f64": (i32, [vp, i32, vp, vp, f64p]),
    "ofh_norm2_f64": (i32, [vp, i32, vp, f64p]),
    "ofh_axpy_f64": (i32, [vp, i32, f64, vp, vp]),
    "ofh_prec_apply": (i32, [vp, i32, vp, vp]),
    "ofh_prec_trans_apply": (i32, [vp, i32, vp, vp]),
    "ofh_ssor_solve": (i32, [vp, vp, vp]),
    "ofh_cg": (i32, [vp, i32, vp, vp, i32, f64, vp]),
    "ofh_cg_noalloc": (i32, [vp, i32, vp, vp, i32, f64, vp]),
    "ofh_ctx_profile": (i32, [vp, i32]),
    "ofh_ctx_profile_read": (i32, [vp, i32, f64p, i32, C.POINTER(i32)]),
    "ofh_bicgstab": (i32, [vp, i32, vp, vp, i32, f64, vp]),
    "ofh_cgs": (i32, [vp, i32, vp, vp, i32, f64, vp]),
    "ofh_bicg": (i32, [vp, i32, vp, vp, i32, f64, vp]),
    "ofh_gmres": (i32, [vp, i32, vp, vp, i32, i32, f64, vp]),
    "ofh_linear_solve": (i32, [vp, i32, i32, vp, vp, i32, f64, vp]),
    "ofh_cg_mixed": (i32, [vp, i32, vp, vp, i32, f64, vp]),
    "ofh_batch_create": (i32, [vp, i32, i32p, C.POINTER(vp)]),
    "ofh_batch_destroy": (i32, [vp]),
    "ofh_batch_cg": (i32, [vp, i32, i32, vp, vp, i32, f64, i32p, i32p, f64p, vp]),
    "ofh_batch_use_graph": (i32, [vp, i32]),
    "ofh_batch_set_config": (i32, [vp, i32, i32]),
    "ofh_batch_get_config": (i32, [vp, C.POINTER(i32), C.POINTER(i32)]),
    "ofh_batch_autotune": (i32, [vp, i32, i32, vp, vp, i32, f64, C.POINTER(i32), C.POINTER(i32),
                                 f64p]),
    "ofh_mesh2d_structured": (i32, [f64, f64, f64, f64, i32, i32, i32, i32, i32, i32, i32, i32,
                                    f64p, i32p, i32p]),
    "ofh_mesh3d_kuhn": (i32, [f64] * 6 + [i32] * 9 + [f64p, i32p, i32p]),
    "ofh_mesh_read": (i32, [C.c_char_p, C.POINTER(vp)]),
    "ofh_mesh_info": (i32, [vp] + [C.POINTER(i32)] * 10),
    "ofh_mesh_get": (i32, [vp, f64p, i64, i32p, i64, i32p, i64, i32p, i64, i32p, i64, i32p, i64]),
    "ofh_mesh_material": (i32, [vp, i32, C.POINTER(i32), C.c_char_p, i32]),
    "ofh_mesh_destroy": (i32, [vp]),
    "ofh_mesh_write": (i32, [C.c_char_p, i32, i32, i32, f64p, i32p, i32, i32, i32, i32p, i32p,
                             i32, i32, i32, i32p, i32p, i32, i32p, C.POINTER(C.c_char_p)]),
    "ofh_number_equations": (i32, [i32, i32, i32p, i32p, C.POINTER(i32)]),
    "ofh_gauss": (i32, [i32, f64p, f64p]),
    "ofh_asm_create": (i32, [vp, i32, i32, f64p, i32, i32, i32p, i32, i32p, C.POINTER(vp)]),
    "ofh_asm_destroy": (i32, [vp]),
    "ofh_asm_info": (i32, [vp, C.POINTER(i32), C.POINTER(i64)]),
    "ofh_asm_matrix": (i32, [vp, C.POINTER(vp)]),
    "ofh_asm_get_dof": (i32, [vp, i32p]),
    "ofh_asm_build": (i32, [vp, i32, vp, vp, vp, vp, f64p, i32, i32, vp]),
    "ofh_asm_traction": (i32, [vp, i32, i32p, vp, vp]),
    "ofh_asm_insert_bc": (i32, [vp, vp, vp, vp]),
    "ofh_proj_create": (i32, [vp, i32, i32, C.POINTER(vp)]),
    "ofh_proj_destroy": (i32, [vp]),
    "ofh_proj_reset": (i32, [vp]),
    "ofh_proj_size": (i32, [vp, C.POINTER(i32)]),
    "ofh_proj_guess": (i32, [vp, vp, vp]),
    "ofh_proj_update": (i32, [vp, vp]),
    "ofh_mpnn_create": (i32, [vp, i32, i32, i32, C.c_uint64, C.POINTER(vp)]),
    "ofh_mpnn_destroy": (i32, [vp]),
    "ofh_mpnn_refresh": (i32, [vp]),
    "ofh_mpnn_num_params": (i32, [vp, C.POINTER(i64)]),
    "ofh_mpnn_get_params": (i32, [vp, f32p]),
    "ofh_mpnn_set_params": (i32, [vp, f32p]),
    "ofh_mpnn_use_graph": (i32, [vp, i32]),
    "ofh_mpnn_set_objective": (i32, [vp, i32]),
    "ofh_mpnn_guess": (i32, [vp, vp, vp, vp, vp]),
    "ofh_mpnn_loss_grad": (i32, [vp, vp, vp, vp, vp, vp, f64p, f32p]),
    "ofh_mpnn_train_step": (i32, [vp, vp, vp, vp, vp, vp, C.c_float, f64p]),
}


def declare(L):
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(L, name)
        fn.restype = res
        fn.argtypes = args
