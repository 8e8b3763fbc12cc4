"""
Enzyme reverse rule of the device stout map `stout_smear!(out, U, nn, links)` (Enzyme cannot
differentiate through the ccall): the forward pass runs the device smearing, the reverse pass
adds the pullback of `out`'s shadow (gf_stout_backward, through `GaugefieldsHIP.stout_pullback`)
to the shadow of `links` and clears `out`'s.  This is the custom-rule seam the reference's
Enzyme extension uses for its lattice products (ext/GaugefieldsEnzymeExt.jl:115-150, the rules
of docs/src/autodiff.md:39-50).
"""
module GaugefieldsHIPEnzymeExt

using Enzyme
import Enzyme.EnzymeRules
import GaugefieldsHIP: stout_smear!, stout_pullback

function EnzymeRules.augmented_primal(config::EnzymeRules.RevConfigWidth{1},
                                      func::Const{typeof(stout_smear!)}, ::Type{<:Const},
                                      out::Duplicated, U::Const, nn::Const, links::Duplicated)
    stout_smear!(out.val, U.val, nn.val, links.val)
    # keep the input if the caller may overwrite it before the reverse pass
    tape = EnzymeRules.overwritten(config)[5] ? deepcopy(links.val) : nothing
    return EnzymeRules.AugmentedReturn(nothing, nothing, tape)
end

function EnzymeRules.reverse(config::EnzymeRules.RevConfigWidth{1},
                             func::Const{typeof(stout_smear!)}, ::Type{<:Const}, tape,
                             out::Duplicated, U::Const, nn::Const, links::Duplicated)
    x = tape === nothing ? links.val : tape
    Ū = stout_pullback(U.val, nn.val, x, out.dval)
    for μ in 1:4
        links.dval[μ] .+= Ū[μ]
        fill!(out.dval[μ], 0)
    end
    return (nothing, nothing, nothing, nothing)
end

end # module
